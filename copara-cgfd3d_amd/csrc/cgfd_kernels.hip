// cgfd_kernels.hip -- gfx950 kernels of the CGFD3D time-stepping hot path.
//
// One RK stage of the reference (RKite, calculate.cu:1355-1599) runs as two
// launches:
//   timg_kernel   the traction image of the top three rows (CalTIMG,
//                 calculate.cu:5428-5666).  It reads only the stage input,
//                 so it runs first and parks its momentum rates in a small
//                 buffer;
//   stage_kernel  everything else, fused: the 27 one-sided derivatives
//                 (CalDiffCL 4623-4801), the curvilinear contraction with
//                 Hooke's law and the momentum equation (CalWaveCL 4804-5118),
//                 the ADE CFS-PML auxiliary rates, the traction-image override
//                 of the top rows, the point sources (LoadForce 6130-6191,
//                 LoadMoment 6193-6255) and the RK4 update of the wavefield
//                 and of the auxiliary fields (Iteration*, 6330-6852),
//                 including the peak-ground-velocity maximum of the last
//                 stage.  The reference's 27 derivative arrays, its rate
//                 arrays and its per-stage state copies never exist.
//
// stage_kernel mapping: a 64-lane wavefront owns 64 consecutive x of one
// (y) row; a workgroup stacks `rows` such rows; blockIdx.z picks a chunk of
// z planes that each lane sweeps upward.  The z-direction stencil taps live
// in a 5-deep register window per component, so each plane of the stage
// input is fetched from memory once by its own lane; x and y taps are read
// from neighbouring lanes' lines through the vector L1/L2.
#include "cgfd_kernels.h"

namespace cgfd {

// ---------------------------------------------------------------------------
// fused RK stage
// ---------------------------------------------------------------------------

// Loads and stores through a uniform field base plus a 32-bit byte offset, so
// the compiler can use the SGPR-base addressing form (one 32-bit VGPR per
// address instead of 64-bit pointer arithmetic).  Offsets never exceed one
// field (< 4 GiB).
template <typename R> __device__ __forceinline__ R ldg(const R *base, unsigned off)
{
    return *(const R *)((const char *)base + off);
}
template <typename R> __device__ __forceinline__ void stg(R *base, unsigned off, R v)
{
    *(R *)((char *)base + off) = v;
}

// Streams touched once per stage (media, metrics, the RK operands FW and tW,
// every store) carry the nontemporal hint, so that the L2 keeps the stage
// input planes, whose lines neighbouring tiles re-read as stencil halos.
#ifndef CGFD_NT
#define CGFD_NT 1
#endif
// FP32 single-use streams go through buffer instructions: the field base
// becomes a resource descriptor formed on the scalar unit, and every load or
// store of a cell shares the one 32-bit lane offset, so no per-lane 64-bit
// address is formed.  The descriptor's record count is the full 32-bit
// range, which no offset here reaches (offsets stay inside one field).
#ifndef CGFD_BUFFER_OPS
#define CGFD_BUFFER_OPS 0
#endif
constexpr int kBufNT = 2; // cache policy of the buffer instructions: nontemporal
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void *base)
{
    return __builtin_amdgcn_make_buffer_rsrc((void *)base, (short)0, -1, 0x00020000);
}
template <typename R> __device__ __forceinline__ R ldg_once(const R *base, unsigned off)
{
#if CGFD_NT
#if CGFD_BUFFER_OPS
    if constexpr (sizeof(R) == 4)
        return __builtin_bit_cast(R, __builtin_amdgcn_raw_buffer_load_b32(buf_rsrc(base), off, 0, kBufNT));
#endif
    return __builtin_nontemporal_load((const R *)((const char *)base + off));
#else
    return ldg(base, off);
#endif
}
template <typename R> __device__ __forceinline__ void stg_once(R *base, unsigned off, R v)
{
#if CGFD_NT
#if CGFD_BUFFER_OPS
    if constexpr (sizeof(R) == 4) {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), buf_rsrc(base), off, 0, kBufNT);
        return;
    }
#endif
    __builtin_nontemporal_store(v, (R *)((char *)base + off));
#else
    stg(base, off, v);
#endif
}
template <typename R> __device__ __forceinline__ R ld_once(const R *p)
{
#if CGFD_NT
    return __builtin_nontemporal_load(p);
#else
    return *p;
#endif
}
template <typename R> __device__ __forceinline__ void st_once(R *p, R v)
{
#if CGFD_NT
    __builtin_nontemporal_store(v, p);
#else
    *p = v;
#endif
}

// member-wise copy of the geometry out of a constant-address-space block
template <typename AP> __device__ __forceinline__ Geom geom_of(AP a)
{
    Geom g;
    g.ni = a->g.ni;
    g.nj = a->g.nj;
    g.nk = a->g.nk;
    g.nxp = a->g.nxp;
    g.nyp = a->g.nyp;
    g.plane = a->g.plane;
    g.fs = a->g.fs;
    g.base = a->g.base;
    return g;
}

// per-(x,y)-column constants of the PML and source bookkeeping
template <typename R> struct Column {
    R APDx, Bx, DBx, APDy, By, DBy;
    bool inx, iny;
    int ix, iy;
    unsigned cxy; // element offset of (x,y) inside one [nj][nxp] map
    int col;      // y*ni + x, the source-table column id
};

template <typename R, typename AP> __device__ __forceinline__ Column<R> column_of(AP a, int x, int y)
{
    const Geom g = geom_of(a);
    Column<R> k;
    const R *pmx = a->pml;
    const R *pmy = a->pml + 3 * g.ni;
    k.APDx = pmx[x];
    k.Bx = pmx[g.ni + x];
    k.DBx = pmx[2 * g.ni + x];
    k.APDy = pmy[y];
    k.By = pmy[g.nj + y];
    k.DBy = pmy[2 * g.nj + y];
    k.inx = x < a->nabs[0] || x >= g.ni - a->nabs[1];
    k.ix = x < a->nabs[0] ? x : x - (g.ni - a->nabs[1]) + a->nabs[0];
    k.iny = y < a->nabs[2] || y >= g.nj - a->nabs[3];
    k.iy = y < a->nabs[2] ? y : y - (g.nj - a->nabs[3]) + a->nabs[2];
    k.cxy = (unsigned)(y * g.nxp + x);
    k.col = y * g.ni + x;
    return k;
}

// first source entry of this column at or above z0 (entries are sorted by
// column, then z)
template <typename AP> __device__ __forceinline__ int first_source(AP a, int col, int z0)
{
    int e = a->col_first ? a->col_first[col] : -1;
    if (e >= 0)
        while (a->sent[e].w == col && a->sent[e].x < z0) ++e;
    return e;
}

// Per-cell inputs that do not depend on the stencil: the media, the nine
// metrics and the RK operands (FW in stages 1-2, tW in stages 1-3).  They are
// fetched at the top of each plane so their latency overlaps the stencil
// arithmetic instead of stalling the RK update.
template <typename R> struct CellIn {
    R lam2mu, miu, rhoinv;
    R m[9];
    R fw[NF], tw[NF];
};

// HM = 1: the horizontal metrics are known to be xi_x = eta_y = 1 and
// xi_y = xi_z = eta_x = eta_z = 0 in every valid cell (grids with fixed x and
// y spacing, vertically stretched: the reference's vmap meshes,
// gridmesh.cpp:101-129, 386-393).  Only the three zeta metrics are loaded;
// the others enter every product as exact constants, so each operation
// rounds exactly as with the loaded values.  A loaded -0 becomes +0, which
// can flip the sign of an exactly-zero intermediate but no value: nothing on
// this path divides by, or otherwise distinguishes, a signed zero.
template <typename R, int ST, int HM = 0, typename AP>
__device__ __forceinline__ void load_cell(AP a, unsigned pb, long fs, CellIn<R> &in)
{
    in.lam2mu = ldg_once(a->med, pb);
    in.miu = ldg_once(a->med + fs, pb);
    in.rhoinv = ldg_once(a->med + 2 * fs, pb);
    if (HM) {
        in.m[XIX] = (R)1;
        in.m[XIY] = (R)0;
        in.m[XIZ] = (R)0;
        in.m[ETX] = (R)0;
        in.m[ETY] = (R)1;
        in.m[ETZ] = (R)0;
#pragma unroll
        for (int m = ZTX; m <= ZTZ; ++m) in.m[m] = ldg_once(a->met + m * fs, pb);
    } else {
#pragma unroll
        for (int m = 0; m < 9; ++m) in.m[m] = ldg_once(a->met + m * fs, pb);
    }
#pragma unroll
    for (int c = 0; c < NF; ++c) {
        if (ST == 1 || ST == 2) in.fw[c] = ldg_once(a->fw + c * fs, pb);
        if (ST >= 1) in.tw[c] = ldg_once((const R *)a->tw + c * fs, pb);
    }
}

// ADE auxiliary operands of one PML cell on one axis, all fetched before any
// of the cell's auxiliary stores (the store targets may alias as far as the
// compiler knows, so loads interleaved with stores would each wait for the
// memory round trip).
template <typename R> struct AuxIn {
    R ai[NF]; // stage input A
    R af[NF]; // step-start FA (stages 1-2)
    R at[NF]; // accumulator tA (stages 1-3)
};

template <typename R, int ST, typename AP>
__device__ __forceinline__ void load_aux(AP a, int ax, long aidx, AuxIn<R> &v)
{
    const long as = a->astr[ax];
    const unsigned off = (unsigned)(aidx * (long)sizeof(R)); // one auxiliary component is far below 4 GiB
#pragma unroll
    for (int c = 0; c < NF; ++c) {
        v.ai[c] = ldg_once(a->ain[ax] + c * as, off);
        if (ST == 1 || ST == 2) v.af[c] = ldg_once(a->af[ax] + c * as, off);
        if (ST >= 1) v.at[c] = ldg_once((const R *)a->at[ax] + c * as, off);
    }
}

// RK update of the auxiliary fields (IterationA*, calculate.cu:6330-6852)
template <typename R, int ST, typename AP>
__device__ __forceinline__ void store_aux(AP a, int ax, long aidx, const AuxIn<R> &v, const R (&ha)[NF])
{
    const long as = a->astr[ax];
    const unsigned off = (unsigned)(aidx * (long)sizeof(R));
    R *ao = a->aout[ax];
    R *at = a->at[ax];
#pragma unroll
    for (int c = 0; c < NF; ++c) {
        if (ST == 0) {
            stg_once(ao + c * as, off, v.ai[c] + a->alpha * ha[c]);
            stg_once(at + c * as, off, v.ai[c] + a->beta * ha[c]);
        } else if (ST < 3) {
            stg_once(ao + c * as, off, v.af[c] + a->alpha * ha[c]);
            stg_once(at + c * as, off, v.at[c] + a->beta * ha[c]);
        } else {
            stg_once(ao + c * as, off, v.at[c] + a->beta * ha[c]);
        }
    }
}

// Everything of one cell after its 27 derivatives are known, up to its nine
// rates hv: the free-surface velocity closure, CalWaveCL, the ADE CFS-PML
// rates with their RK update, the traction-image override and the point
// sources.  The RK update of the wavefield is rk_store (below).
//   w: z window of the stage input (slot CI is the cell itself)
template <typename R, int SX, int SZ, int ST, int HY, typename AP>
__device__ __forceinline__ void cell_rates(AP a, const Column<R> &k, const CellIn<R> &in, R (&Dx)[NF], R (&Dy)[NF],
                                           R (&Dz)[NF], const R (&w)[NF][5], int x, int y, int z, int &e,
                                           int tindex, R (&hv)[NF])
{
    const Geom g = geom_of(a);
    const double h = a->h;
    constexpr int CI = SZ > 0 ? 1 : 3;

    // velocity zeta-derivatives of the top three rows (CondFreeVLOW,
    // calculate.cu:4736-4769): closure on the surface, 2-point one row
    // below, 2-4 MacCormack two rows below
    const int top = z - (g.nk - 3);
    R cx[9], cy[9];
    if (top == 2) {
#pragma unroll
        for (int m = 0; m < 9; ++m) {
            cx[m] = a->cov[m * a->cstr + k.cxy];
            cy[m] = a->cov[(9 + m) * a->cstr + k.cxy];
        }
        Dz[VX] = cx[0] * Dx[VX] + cx[1] * Dx[VY] + cx[2] * Dx[VZ] + cy[0] * Dy[VX] + cy[1] * Dy[VY] + cy[2] * Dy[VZ];
        Dz[VY] = cx[3] * Dx[VX] + cx[4] * Dx[VY] + cx[5] * Dx[VZ] + cy[3] * Dy[VX] + cy[4] * Dy[VY] + cy[5] * Dy[VZ];
        Dz[VZ] = cx[6] * Dx[VX] + cx[7] * Dx[VY] + cx[8] * Dx[VZ] + cy[6] * Dy[VX] + cy[7] * Dy[VY] + cy[8] * Dy[VZ];
    } else if (top == 1) {
#pragma unroll
        for (int c = VX; c <= VZ; ++c)
            Dz[c] = SZ > 0 ? (w[c][2] - w[c][1]) / a->hr : (w[c][3] - w[c][2]) / a->hr;
    } else if (top == 0) {
#pragma unroll
        for (int c = VX; c <= VZ; ++c) {
            double s = SZ > 0 ? M24A0 * (double)w[c][1] + M24A1 * (double)w[c][2] + M24A2 * (double)w[c][3]
                              : M24A2 * (double)w[c][1] + M24A1 * (double)w[c][2] + M24A0 * (double)w[c][3];
            Dz[c] = (R)(SZ > 0 ? s / h : s / (-h));
        }
    }

    // ---- CalWaveCL: contraction with metrics, Hooke, momentum -------------
    // media are stored as lam2mu = rho*alpha*alpha, miu = rho*beta*beta and
    // 1/rho, each rounded exactly as calculate.cu:4855-4858 computes them
    const R lam2mu = in.lam2mu, miu = in.miu, rho = in.rhoinv;
    const R lambda = lam2mu - (R)2 * miu; // 2*miu is exact: one rounding, as the reference's double expression
    const R xix = in.m[XIX], xiy = in.m[XIY], xiz = in.m[XIZ];
    const R etx = in.m[ETX], ety = in.m[ETY], etz = in.m[ETZ];
    const R ztx = in.m[ZTX], zty = in.m[ZTY], ztz = in.m[ZTZ];

    // Contraction with the metrics.  With the hybrid grid (hybrid_grid = 1),
    // planes below the conversion depth (z < zcon, uniform over the
    // workgroup) take CalWave's Cartesian form, which uses only xi_x, eta_y
    // and zeta_z in its own operand order (calculate.cu:3910-3956);
    // elsewhere the curvilinear contraction of CalWaveCL (4855-4930).  HY = 0
    // compiles the curvilinear form alone (grids without a Cartesian part).
    R DxiVx, DetVx, DztVx, DxiVy, DetVy, DztVy, DxiVz, DetVz, DztVz;
    R DxiTxx, DetTxx, DztTxx, DxiTyy, DetTyy, DztTyy, DxiTzz, DetTzz, DztTzz;
    R DxiTxy, DetTxy, DztTxy, DxiTxz, DetTxz, DztTxz, DxiTyz, DetTyz, DztTyz;
    if (!HY || z >= a->zcon) {
        DxiVx = (xix * Dx[TXX] + xiy * Dx[TXY] + xiz * Dx[TXZ]) * rho;
        DetVx = (etx * Dy[TXX] + ety * Dy[TXY] + etz * Dy[TXZ]) * rho;
        DztVx = (ztx * Dz[TXX] + zty * Dz[TXY] + ztz * Dz[TXZ]) * rho;
        DxiVy = (xix * Dx[TXY] + xiy * Dx[TYY] + xiz * Dx[TYZ]) * rho;
        DetVy = (etx * Dy[TXY] + ety * Dy[TYY] + etz * Dy[TYZ]) * rho;
        DztVy = (ztx * Dz[TXY] + zty * Dz[TYY] + ztz * Dz[TYZ]) * rho;
        DxiVz = (xix * Dx[TXZ] + xiy * Dx[TYZ] + xiz * Dx[TZZ]) * rho;
        DetVz = (etx * Dy[TXZ] + ety * Dy[TYZ] + etz * Dy[TZZ]) * rho;
        DztVz = (ztx * Dz[TXZ] + zty * Dz[TYZ] + ztz * Dz[TZZ]) * rho;

        DxiTxx = lam2mu * xix * Dx[VX] + lambda * xiy * Dx[VY] + lambda * xiz * Dx[VZ];
        DetTxx = lam2mu * etx * Dy[VX] + lambda * ety * Dy[VY] + lambda * etz * Dy[VZ];
        DztTxx = lam2mu * ztx * Dz[VX] + lambda * zty * Dz[VY] + lambda * ztz * Dz[VZ];
        DxiTyy = lambda * xix * Dx[VX] + lam2mu * xiy * Dx[VY] + lambda * xiz * Dx[VZ];
        DetTyy = lambda * etx * Dy[VX] + lam2mu * ety * Dy[VY] + lambda * etz * Dy[VZ];
        DztTyy = lambda * ztx * Dz[VX] + lam2mu * zty * Dz[VY] + lambda * ztz * Dz[VZ];
        DxiTzz = lambda * xix * Dx[VX] + lambda * xiy * Dx[VY] + lam2mu * xiz * Dx[VZ];
        DetTzz = lambda * etx * Dy[VX] + lambda * ety * Dy[VY] + lam2mu * etz * Dy[VZ];
        DztTzz = lambda * ztx * Dz[VX] + lambda * zty * Dz[VY] + lam2mu * ztz * Dz[VZ];
        DxiTxy = (xiy * Dx[VX] + xix * Dx[VY]) * miu;
        DetTxy = (ety * Dy[VX] + etx * Dy[VY]) * miu;
        DztTxy = (zty * Dz[VX] + ztx * Dz[VY]) * miu;
        DxiTxz = (xiz * Dx[VX] + xix * Dx[VZ]) * miu;
        DetTxz = (etz * Dy[VX] + etx * Dy[VZ]) * miu;
        DztTxz = (ztz * Dz[VX] + ztx * Dz[VZ]) * miu;
        DxiTyz = (xiz * Dx[VY] + xiy * Dx[VZ]) * miu;
        DetTyz = (etz * Dy[VY] + ety * Dy[VZ]) * miu;
        DztTyz = (ztz * Dz[VY] + zty * Dz[VZ]) * miu;
    } else {
        DxiVx = rho * xix * Dx[TXX];
        DetVx = rho * ety * Dy[TXY];
        DztVx = ztz * Dz[TXZ] * rho;
        DxiVy = rho * xix * Dx[TXY];
        DetVy = rho * ety * Dy[TYY];
        DztVy = ztz * Dz[TYZ] * rho;
        DxiVz = rho * xix * Dx[TXZ];
        DetVz = rho * ety * Dy[TYZ];
        DztVz = ztz * Dz[TZZ] * rho;
        DxiTxx = lam2mu * xix * Dx[VX];
        DetTxx = lambda * ety * Dy[VY];
        DztTxx = lambda * ztz * Dz[VZ];
        DxiTyy = lambda * xix * Dx[VX];
        DetTyy = lam2mu * ety * Dy[VY];
        DztTyy = lambda * ztz * Dz[VZ];
        DxiTzz = lambda * xix * Dx[VX];
        DetTzz = lambda * ety * Dy[VY];
        DztTzz = lam2mu * ztz * Dz[VZ];
        DxiTxy = miu * xix * Dx[VY];
        DetTxy = miu * ety * Dy[VX];
        DztTxy = (R)0;
        DxiTxz = miu * xix * Dx[VZ];
        DetTxz = (R)0;
        DztTxz = ztz * Dz[VX] * miu;
        DxiTyz = (R)0;
        DetTyz = miu * ety * Dy[VZ];
        DztTyz = ztz * Dz[VY] * miu;
    }

    const R *pmz = a->pml + 3 * g.ni + 3 * g.nj;
    const R APDz = pmz[z], Bz = pmz[g.nk + z], DBz = pmz[2 * g.nk + z];
    const R Bx = k.Bx, By = k.By;
    // CalWaveCL divides each axis' term by that axis' CFS factor B.  x/1 is
    // x exactly, so an axis whose B is 1 for every lane of the wavefront
    // (all of the interior; By is uniform over a wavefront row and Bz over a
    // plane) skips its nine divisions; the sums keep the reference's order.
    R qx[NF] = {DxiVx, DxiVy, DxiVz, DxiTxx, DxiTyy, DxiTzz, DxiTxy, DxiTxz, DxiTyz};
    R qy[NF] = {DetVx, DetVy, DetVz, DetTxx, DetTyy, DetTzz, DetTxy, DetTxz, DetTyz};
    R qz[NF] = {DztVx, DztVy, DztVz, DztTxx, DztTyy, DztTzz, DztTxy, DztTxz, DztTyz};
    if (Bx != (R)1) {
#pragma unroll
        for (int c = 0; c < NF; ++c) qx[c] = qx[c] / Bx;
    }
    if (By != (R)1) {
#pragma unroll
        for (int c = 0; c < NF; ++c) qy[c] = qy[c] / By;
    }
    if (Bz != (R)1) {
#pragma unroll
        for (int c = 0; c < NF; ++c) qz[c] = qz[c] / Bz;
    }
#pragma unroll
    for (int c = 0; c < NF; ++c) hv[c] = qx[c] + qy[c] + qz[c];

    const unsigned tcell = (unsigned)(top > 0 ? top : 0) * (unsigned)(g.nj * g.nxp) + k.cxy; // top-row slot

    // ---- ADE CFS-PML (calculate.cu:4993-5107) + auxiliary RK update --------
    // The rates of one axis feed that axis' RK update straight away; the
    // traction image replaces the velocity rates of the top rows first.
    R DzVx1 = 0, DzVx2 = 0, DzVy1 = 0, DzVy2 = 0, DzVz1 = 0, DzVz2 = 0;
    if (top == 2 && (k.inx || k.iny)) {
        // surface contributions routed into the x/y auxiliary stress rates
        DzVx1 = cx[0] * Dx[VX] + cx[1] * Dx[VY] + cx[2] * Dx[VZ];
        DzVx2 = cy[0] * Dy[VX] + cy[1] * Dy[VY] + cy[2] * Dy[VZ];
        DzVy1 = cx[3] * Dx[VX] + cx[4] * Dx[VY] + cx[5] * Dx[VZ];
        DzVy2 = cy[3] * Dy[VX] + cy[4] * Dy[VY] + cy[5] * Dy[VZ];
        DzVz1 = cx[6] * Dx[VX] + cx[7] * Dx[VY] + cx[8] * Dx[VZ];
        DzVz2 = cy[6] * Dy[VX] + cy[7] * Dy[VY] + cy[8] * Dy[VZ];
    }
    if (k.inx) {
        const R APDx = k.APDx, DBx = k.DBx;
        const long aidx = ((long)z * g.nj + y) * a->npml[0] + k.ix;
        AuxIn<R> v;
        load_aux<R, ST>(a, 0, aidx, v);
        R ha[NF];
#pragma unroll
        for (int c = 0; c < NF; ++c) hv[c] -= v.ai[c] / Bx;
        ha[VX] = DxiVx * DBx - APDx * v.ai[VX];
        ha[VY] = DxiVy * DBx - APDx * v.ai[VY];
        ha[VZ] = DxiVz * DBx - APDx * v.ai[VZ];
        ha[TXX] = DxiTxx * DBx - APDx * v.ai[TXX];
        ha[TYY] = DxiTyy * DBx - APDx * v.ai[TYY];
        ha[TZZ] = DxiTzz * DBx - APDx * v.ai[TZZ];
        ha[TXY] = DxiTxy * DBx - APDx * v.ai[TXY];
        ha[TXZ] = DxiTxz * DBx - APDx * v.ai[TXZ];
        ha[TYZ] = DxiTyz * DBx - APDx * v.ai[TYZ];
        if (top == 2) {
            ha[TXX] += DBx * Bx * (lam2mu * ztx * DzVx1 + lambda * zty * DzVy1 + lambda * ztz * DzVz1);
            ha[TYY] += DBx * Bx * (lambda * ztx * DzVx1 + lam2mu * zty * DzVy1 + lambda * ztz * DzVz1);
            ha[TZZ] += DBx * Bx * (lambda * ztx * DzVx1 + lambda * zty * DzVy1 + lam2mu * ztz * DzVz1);
            ha[TXY] += DBx * Bx * (zty * DzVx1 + ztx * DzVy1) * miu;
            ha[TXZ] += DBx * Bx * (ztz * DzVx1 + ztx * DzVz1) * miu;
            ha[TYZ] += DBx * Bx * (ztz * DzVy1 + zty * DzVz1) * miu;
        }
        if (top >= 0) {
#pragma unroll
            for (int c = VX; c <= VZ; ++c) ha[c] = a->timg[(3 + c) * a->tstr + tcell];
        }
        store_aux<R, ST>(a, 0, aidx, v, ha);
    }
    if (k.iny) {
        const R APDy = k.APDy, DBy = k.DBy;
        const long aidx = ((long)z * a->npml[1] + k.iy) * g.nxp + x;
        AuxIn<R> v;
        load_aux<R, ST>(a, 1, aidx, v);
        R ha[NF];
#pragma unroll
        for (int c = 0; c < NF; ++c) hv[c] -= v.ai[c] / By;
        ha[VX] = DetVx * DBy - APDy * v.ai[VX];
        ha[VY] = DetVy * DBy - APDy * v.ai[VY];
        ha[VZ] = DetVz * DBy - APDy * v.ai[VZ];
        ha[TXX] = DetTxx * DBy - APDy * v.ai[TXX];
        ha[TYY] = DetTyy * DBy - APDy * v.ai[TYY];
        ha[TZZ] = DetTzz * DBy - APDy * v.ai[TZZ];
        ha[TXY] = DetTxy * DBy - APDy * v.ai[TXY];
        ha[TXZ] = DetTxz * DBy - APDy * v.ai[TXZ];
        ha[TYZ] = DetTyz * DBy - APDy * v.ai[TYZ];
        if (top == 2) {
            ha[TXX] += DBy * By * (lam2mu * ztx * DzVx2 + lambda * zty * DzVy2 + lambda * ztz * DzVz2);
            ha[TYY] += DBy * By * (lambda * ztx * DzVx2 + lam2mu * zty * DzVy2 + lambda * ztz * DzVz2);
            ha[TZZ] += DBy * By * (lambda * ztx * DzVx2 + lambda * zty * DzVy2 + lam2mu * ztz * DzVz2);
            ha[TXY] += DBy * By * (zty * DzVx2 + ztx * DzVy2) * miu;
            ha[TXZ] += DBy * By * (ztz * DzVx2 + ztx * DzVz2) * miu;
            ha[TYZ] += DBy * By * (ztz * DzVy2 + zty * DzVz2) * miu;
        }
        if (top >= 0) {
#pragma unroll
            for (int c = VX; c <= VZ; ++c) ha[c] = a->timg[(6 + c) * a->tstr + tcell];
        }
        store_aux<R, ST>(a, 1, aidx, v, ha);
    }
    if (z < a->nabs[4] || z >= g.nk - a->nabs[5]) {
        const int iz = z < a->nabs[4] ? z : z - (g.nk - a->nabs[5]) + a->nabs[4];
        const long aidx = ((long)iz * g.nj + y) * g.nxp + x;
        AuxIn<R> v;
        load_aux<R, ST>(a, 2, aidx, v);
        R ha[NF];
#pragma unroll
        for (int c = 0; c < NF; ++c) hv[c] -= v.ai[c] / Bz;
        ha[VX] = DztVx * DBz - APDz * v.ai[VX];
        ha[VY] = DztVy * DBz - APDz * v.ai[VY];
        ha[VZ] = DztVz * DBz - APDz * v.ai[VZ];
        ha[TXX] = DztTxx * DBz - APDz * v.ai[TXX];
        ha[TYY] = DztTyy * DBz - APDz * v.ai[TYY];
        ha[TZZ] = DztTzz * DBz - APDz * v.ai[TZZ];
        ha[TXY] = DztTxy * DBz - APDz * v.ai[TXY];
        ha[TXZ] = DztTxz * DBz - APDz * v.ai[TXZ];
        ha[TYZ] = DztTyz * DBz - APDz * v.ai[TYZ];
        store_aux<R, ST>(a, 2, aidx, v, ha);
    }

    // ---- CalTIMG override of the velocity rates on the top three rows -----
    if (top >= 0) {
#pragma unroll
        for (int c = VX; c <= VZ; ++c) hv[c] = a->timg[c * a->tstr + tcell];
    }

    // ---- LoadForce, LoadMoment then LoadRmom, in source order ---------------
    if (e >= 0) {
        int4 en = a->sent[e];
        while (en.w == k.col && en.x == z) {
            const R *cf = a->scoef + (long)e * 8;
            if (en.y == 0) {
                const R stf = a->fstf[(long)en.z * a->nstf_f + tindex];
                const R A = cf[0];
                hv[VX] += stf * cf[1] * A;
                hv[VY] += stf * cf[2] * A;
                hv[VZ] += stf * cf[3] * A;
            } else if (en.y == 1) {
                const R stf = a->mstf[(long)en.z * a->nstf_m + tindex];
                R A = (R)1;
                A = stf * A / cf[0];
                hv[TXX] -= cf[1] * A;
                hv[TYY] -= cf[2] * A;
                hv[TZZ] -= cf[3] * A;
                hv[TXY] -= cf[4] * A;
                hv[TXZ] -= cf[5] * A;
                hv[TYZ] -= cf[6] * A;
            } else {
                // focal point: interpolated moment-rate tensor of this stage
                const R *fv = a->rstf + (long)en.z * 6 * a->nstf_r + tindex;
                const R A = cf[0];
#pragma unroll
                for (int c = 0; c < 6; ++c) hv[TXX + c] -= fv[(long)c * a->nstf_r] * A;
            }
            en = a->sent[++e];
        }
    }

}

// store of the RK update: in FP32 a buffer store whose lane offset lies
// past the field (keep = false) is dropped by the hardware, so lanes outside
// the grid store nothing and no branch is needed; FP64 lanes outside the
// grid store the zeros they computed into their own guard cell
template <typename R>
__device__ __forceinline__ void stg_keep(R *base, unsigned off, R v, bool keep, unsigned bytes)
{
#if CGFD_NT
    if constexpr (sizeof(R) == 4) {
        const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc((void *)base, (short)0, (int)bytes, 0x00020000);
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, keep ? off : 0xfffffff0u, 0, kBufNT);
        return;
    }
#endif
    stg_once(base, off, v);
}

// RK4 update of the wavefield (Iteration*, calculate.cu:6330-6852) from the
// cell's rates hv.  The stage-0 input is FW itself (w[c][CI]).
template <typename R, int SZ, int ST, typename AP>
__device__ __forceinline__ void rk_store(AP a, const CellIn<R> &in, const R (&w)[NF][5], const R (&hv)[NF],
                                         unsigned pb, long fs, bool keep = true)
{
    constexpr int CI = SZ > 0 ? 1 : 3;
    const unsigned fb = (unsigned)(fs * (long)sizeof(R)); // bytes of one field (< 4 GiB)
#pragma unroll
    for (int c = 0; c < NF; ++c) {
        if (ST == 0) {
            const R f0 = w[c][CI];
            stg_keep(a->wout + c * fs, pb, f0 + a->alpha * hv[c], keep, fb);
            stg_keep(a->tw + c * fs, pb, f0 + a->beta * hv[c], keep, fb);
        } else if (ST < 3) {
            stg_keep(a->wout + c * fs, pb, in.fw[c] + a->alpha * hv[c], keep, fb);
            stg_keep(a->tw + c * fs, pb, in.tw[c] + a->beta * hv[c], keep, fb);
        } else {
            stg_keep(a->wout + c * fs, pb, in.tw[c] + a->beta * hv[c], keep, fb);
        }
    }
}

// peak ground velocity of the last stage: running max |V| of the surface
// row (IterationFinalPV, calculate.cu:6689-6694), from the stored values
template <typename R, int ST, typename AP>
__device__ __forceinline__ void pgv_update(AP a, const Column<R> &k, const CellIn<R> &in, const R (&hv)[NF], int top)
{
    if (ST == 3 && top == 2 && a->pgv) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const R wn = in.tw[c] + a->beta * hv[c];
            R *pv = a->pgv + c * a->cstr + k.cxy; // PGV maps share the [nj][nxp] slot size of cov
            *pv = rmax(rabs(wn), rabs(*pv));
        }
    }
}

// the whole cell: rates, RK update, PGV (lanes of valid cells only)
template <typename R, int SX, int SZ, int ST, typename AP>
__device__ __forceinline__ void finish_cell(AP a, const Column<R> &k, const CellIn<R> &in, R (&Dx)[NF], R (&Dy)[NF],
                                            R (&Dz)[NF], const R (&w)[NF][5], int x, int y, int z, unsigned pb,
                                            int &e, long fs, int tindex)
{
    R hv[NF];
    cell_rates<R, SX, SZ, ST, 1>(a, k, in, Dx, Dy, Dz, w, x, y, z, e, tindex, hv);
    rk_store<R, SZ, ST>(a, in, w, hv, pb, fs);
    pgv_update<R, ST>(a, k, in, hv, z - (geom_of(a).nk - 3));
}

// Variant "direct": the x and y stencil taps are read straight from memory
// (vector L1/L2); one lane per (x,y) column sweeps its z chunk with the z taps
// in a 5-deep register window.
template <typename R, int SX, int SZ, int ST>
__global__ void __launch_bounds__(256) stage_direct_kernel(const StageArgs<R> *pa, int tinc)
{
    typedef const __attribute__((address_space(4))) StageArgs<R> *AP;
    const AP a = (AP)pa;
    const int tindex = 2 * *a->step + tinc;
    const Geom g = geom_of(a);
    const int x = blockIdx.x * 64 + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= g.ni || y >= g.nj) return;
    const int z0 = blockIdx.z * a->zlen;
    const int z1 = min(g.nk, z0 + a->zlen);
    if (z0 >= z1) return;

    const long fs = g.fs;
    const long row = g.nxp;
    const long pl = g.plane;
    const double h = a->h, rh = a->rh;
    constexpr int ZOFF = SZ > 0 ? -1 : -3; // window slot 0 sits at z+ZOFF
    constexpr int CI = SZ > 0 ? 1 : 3;     // window slot of the cell itself

    const Column<R> k = column_of<R>(a, x, y);
    int e = first_source(a, k.col, z0);

    long p = gidx(g, x, y, z0);
    R w[NF][5];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int c = 0; c < NF; ++c) w[c][j] = a->win[c * fs + p + (ZOFF + j) * pl];

    for (int z = z0; z < z1; ++z, p += pl) {
#pragma unroll
        for (int c = 0; c < NF; ++c) w[c][4] = a->win[c * fs + p + (ZOFF + 4) * pl];

        // ---- CalDiffCL: one-sided derivatives along xi, eta, zeta -------------
        R Dx[NF], Dy[NF], Dz[NF];
#pragma unroll
        for (int c = 0; c < NF; ++c) {
            const R *f = a->win + c * fs + p;
            const R ctr = w[c][CI];
            if (SX > 0) {
                Dx[c] = dfd<R, 1>(f[-1], ctr, f[1], f[2], f[3], h, rh);
                Dy[c] = dfd<R, 1>(f[-row], ctr, f[row], f[2 * row], f[3 * row], h, rh);
            } else {
                Dx[c] = dfd<R, -1>(f[-3], f[-2], f[-1], ctr, f[1], h, rh);
                Dy[c] = dfd<R, -1>(f[-3 * row], f[-2 * row], f[-row], ctr, f[row], h, rh);
            }
            Dz[c] = dfd<R, SZ>(w[c][0], w[c][1], w[c][2], w[c][3], w[c][4], h, rh);
        }
        CellIn<R> cin;
        load_cell<R, ST>(a, (unsigned)(p * sizeof(R)), fs, cin);
        finish_cell<R, SX, SZ, ST>(a, k, cin, Dx, Dy, Dz, w, x, y, z, (unsigned)(p * sizeof(R)), e, fs, tindex);

#pragma unroll
        for (int c = 0; c < NF; ++c) {
            w[c][0] = w[c][1];
            w[c][1] = w[c][2];
            w[c][2] = w[c][3];
            w[c][3] = w[c][4];
        }
    }
}

// Variant "lds": xy-plane staging.  A workgroup of LDS_TY wavefronts owns a
// 64 x LDS_TY (or 32 x 2*LDS_TY) tile of (x,y) columns and sweeps a chunk of
// z planes.  Per plane
// every lane writes its cell of the 9 stage-input components into an LDS tile
// that also holds the one-sided halo (1 cell behind and 3 ahead for forward
// differences, 3 behind and 1 ahead for backward), so all 18 in-plane stencils
// of a cell are LDS reads at immediate offsets from one address.  The z taps
// stay in the per-lane 5-deep register window; the next plane's global loads
// (window lead and halo) are issued before the current plane's arithmetic,
// unconditionally (lanes outside the grid read a clamped in-bounds address
// and discard the value), so no branch splits the load stream.  A plane's
// halo is fetched together with its window lead and parked in an LDS ring
// until the plane is the centre (see below).
#ifndef CGFD_LDS_TY
#define CGFD_LDS_TY 4
#endif
constexpr int LDS_TY = CGFD_LDS_TY;
static_assert(LDS_TY >= 4 && LDS_TY <= 16 && 4 * LDS_TY <= 64, "x-halo lanes must fit one wavefront");
// FP64 keeps 4-row tiles: its one wavefront per SIMD needs the full register file
template <typename R> constexpr int lds_ty() { return sizeof(R) == 4 ? LDS_TY : 4; }
#ifndef CGFD_LDS_WAVES
#define CGFD_LDS_WAVES 2
#endif
#ifndef CGFD_HALO_RING
#define CGFD_HALO_RING 1
#endif
#define CGFD_LDS_ATTR __attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 ? CGFD_LDS_WAVES : 1, 8)))

template <typename R, int SX, int SZ, int ST, int HM, int TW>
__global__ void __launch_bounds__(64 * lds_ty<R>()) CGFD_LDS_ATTR
stage_lds_kernel(const StageArgs<R> *pa, const int *starts, int nctx, int tinc)
{
    // The argument block lives in device memory and is read through the
    // constant address space, so fields are fetched with scalar loads where
    // they are used instead of being preloaded into (and spilled from) SGPRs.
    // A batched launch (nctx > 1) covers several grids: grid c owns
    // workgroups starts[c] .. starts[c+1]-1 and its block lies c *
    // kCtxArgBytes after pa.  The search runs on scalar registers.  With
    // nctx = -n the n grids are identical and interleaved tile by tile.
    typedef const __attribute__((address_space(4))) StageArgs<R> *AP;
    typedef const __attribute__((address_space(4))) int *SP;
    int ctx = 0, nb = gridDim.x, id = blockIdx.x;
    if (nctx < 0) {
        // n identical grids on one medium: workgroups are dealt round-robin
        // to the 8 XCDs, and the n copies of one tile follow each other on
        // one XCD, which reads the metric and media lines they share from
        // its L2.  Tiles past the last multiple of 8 run grid after grid.
        const int n = -nctx;
        nb = gridDim.x / n;
        const int M = (nb >> 3) * 8 * n;
        if (id < M) {
            const int k = id >> 3;
            ctx = k % n;
            id = (k / n) * 8 + (id & 7);
        } else {
            ctx = (id - M) % n;
            id = (nb >> 3) * 8 + (id - M) / n;
        }
    } else if (nctx > 1) {
        const SP st = (SP)starts;
        int lo = 0, hi = nctx - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (st[mid] <= id) lo = mid;
            else hi = mid - 1;
        }
        ctx = lo;
        nb = st[lo + 1] - st[lo];
        id -= st[lo];
    }
    const AP a0 = (AP)((const char *)pa + (long)ctx * kCtxArgBytes);
    const int tindex = 2 * *a0->step + tinc;
    // A workgroup of TY wavefronts owns a tile of TW columns x TR rows: with
    // TW = 64 a wavefront is one tile row, with TW = 32 it holds two (rows
    // whose length leaves a short last 64-column tile waste fewer lanes).
    static_assert(TW == 64 || TW == 32, "tile width");
    constexpr int TY = lds_ty<R>(), HALVES = 64 / TW, TR = TY * HALVES;
    constexpr int HX = TW + 4, HY = TR + 4;
    constexpr int O = SX > 0 ? 1 : 3; // halo cells below the tile in x and in y
    constexpr int ZOFF = SZ > 0 ? -1 : -3;
    constexpr int CI = SZ > 0 ? 1 : 3;
    // The window lead of plane z is plane z+D.  With the halo ring, the halo
    // rows and columns of a plane are fetched in the same iteration as its
    // lead, i.e. when the neighbouring tiles fetch those same lines as their
    // own lead (so they are still in L2), and wait in an LDS ring of D+1
    // plane tiles until the plane becomes the centre.
    constexpr int D = ZOFF + 4;
    constexpr int HD = CGFD_HALO_RING ? D : 0; // halo prefetch distance
    constexpr int NSLOT = HD + 1;              // 4, 2 or 1 plane tiles
    typedef R LT;                              // type of the plane tiles
    __shared__ LT lds[NSLOT][NF][HY][HX];
    auto slot = [](int p) { return p & (NSLOT - 1); };

    const AP a = a0;
    const Geom g = geom_of(a);
    const int lane = threadIdx.x, wv = threadIdx.y;
    // which tile row of its wavefront the lane is in (blockDim.x is 64)
    const int part = TW == 64 ? 0 : lane / TW;
    const int tx = TW == 64 ? lane : lane % TW, ty = wv * HALVES + part;

    // Tile numbering.  Workgroups are dealt round-robin to the 8 XCDs, so
    // the identity map (default) gives every XCD a uniform sample of the
    // domain (the same share of PML, surface and source work) and keeps the
    // running workgroups inside one z chunk.  Map 0 gives each XCD a
    // contiguous run of tiles (neighbours share its L2, but the XCDs holding
    // the bottom PML planes finish last); map 4 does that within groups of
    // 8*CGFD_TILE_GROUP consecutive tiles.  Both cut the L2 fetch traffic by
    // ~8 % on the 600x600x300 model yet run slower than the identity map.
    const int ntx = (g.ni + TW - 1) / TW, nty = (g.nj + TR - 1) / TR;
#ifndef CGFD_TILE_MAP
#define CGFD_TILE_MAP 1
#endif
#if CGFD_TILE_MAP == 1
    const int L = id;
#elif CGFD_TILE_MAP == 5
    // whole rows of x tiles on one XCD (x neighbours share its L2), rows
    // dealt round-robin to the XCDs; the rows past the last multiple of 8
    // keep the identity map
    const int nrow = nb / ntx, M = (nrow / 8) * 8 * ntx;
    int L = id;
    if (id < M) {
        const int k = id / 8;
        L = ((k / ntx) * 8 + id % 8) * ntx + k % ntx;
    }
#elif CGFD_TILE_MAP == 4
    // groups of 8*S consecutive tiles (about what the chip runs at once);
    // inside a group each XCD takes a contiguous run
#ifndef CGFD_TILE_GROUP
#define CGFD_TILE_GROUP 64
#endif
    constexpr int G = 8 * CGFD_TILE_GROUP;
    const int gbase = id / G * G, rem = min(G, nb - gbase), wi = id - gbase;
    const int q = rem / 8, r = rem % 8, xcd = wi % 8;
    const int L = gbase + xcd * q + min(xcd, r) + wi / 8;
#else
    const int q = nb / 8, r = nb % 8, xcd = id % 8;
    const int L = xcd * q + min(xcd, r) + id / 8;
#endif
    const int bx = L % ntx, by = (L / ntx) % nty, bz = L / (ntx * nty);
    const int x0 = bx * TW, y0 = by * TR;
    const int x = x0 + tx, y = y0 + ty;
    const bool active = x < g.ni && y < g.nj;
    const int z0 = bz * a->zlen;
    const int z1 = min(g.nk, z0 + a->zlen);
    if (z0 >= z1) return; // uniform over the workgroup

    const long fs0 = g.fs;
    const double h = a->h, rh = a->rh;
    const unsigned esz = sizeof(R);
    const unsigned plb = (unsigned)(g.plane * esz); // bytes per plane

    // Addresses.  x < ntx*TW <= nxp always; rows past the last guard row
    // (y >= nyp, only in a ragged last tile) are clamped onto it (a zero row),
    // so every load stays inside its own plane.
    const int yc = min(y, g.nyp - 1);
    // Halo slots of this lane.  Halo rows: wavefront hk (the first 4) stages
    // halo row hk (TW columns); with two tile rows per wavefront its lanes of
    // part 0 take fields 0-4 of that row and those of part 1 fields 5-8, so
    // a lane stages NR fields.  Halo columns: the 4*TR slots of one field
    // (4 columns x TR rows) go to lanes q = lane % TW < 4*TR of one part;
    // wavefront hk, part p stages fields hk + TY*(HALVES*j + p), j < NH.
    // Slots past the last field reload field 8 and discard it, so every
    // lane issues the same loads.
    const int hk = wv;
    const int hrow = SX > 0 ? (hk == 0 ? 0 : TR + hk) : (hk < 3 ? hk : TR + 3);
    const int hy = y0 - O + hrow, hxx = x0 + tx;
    const bool hv1 = hk < 4 && hy >= -3 && hy < g.nyp;
    constexpr int NR = (NF + HALVES - 1) / HALVES;
    // field of halo-row slot j; a plain expression rather than a lambda,
    // whose call inside a field address kept the 64-column kernel's field
    // bases out of scalar registers (36 more 64-bit vector adds)
#define CGFD_HRFIELD(j) (TW == 64 ? (j) : min(part * NR + (j), NF - 1))
    auto hrkeep = [&](int j) { return hk < 4 && part * NR + j < NF; };
    constexpr int NH = (NF + TY * HALVES - 1) / (TY * HALVES);
    const bool xlane = tx < 4 * TR;
    const int hc2k = tx / TR, hr2 = tx % TR;
    const int hcol = SX > 0 ? (hc2k == 0 ? 0 : TW + hc2k) : (hc2k < 3 ? hc2k : TW + 3);
    const int hx2 = x0 - O + hcol, hy2 = y0 + hr2;
    const bool hv2 = xlane && hx2 >= -3 && hx2 < g.nxp && hy2 < g.nyp;
    auto hfield = [&](int j) { return min(hk + TY * (HALVES * j + part), NF - 1); };
    auto hkeep = [&](int j) { return xlane && hk + TY * (HALVES * j + part) < NF; };
    // byte offsets inside plane 0 (may be negative: added to a plane base);
    // invalid halo slots fall back to the lane's own cell
    const int rel_c = (yc * g.nxp + x) * (int)esz;
    const int rel_1 = hv1 ? (hy * g.nxp + hxx) * (int)esz : rel_c;
    const int rel_2 = hv2 ? (hy2 * g.nxp + hx2) * (int)esz : rel_c;
    const unsigned base0 = (unsigned)(g.base * esz);
    auto pbase = [&](int zz) { return base0 + (unsigned)zz * plb; };

    const Column<R> k = column_of<R>(a, active ? x : 0, active ? y : 0);
    int e = active ? first_source(a, k.col, z0) : -1;

    R w[NF][5];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int c = 0; c < NF; ++c) w[c][j] = ldg(a->win + c * fs0, pbase(z0 + ZOFF + j) + rel_c);

    // halos of the planes z0 .. z0+HD-1 go straight into their ring tiles
    for (int p = z0; p < z0 + HD; ++p) {
#pragma unroll
        for (int j = 0; j < NR; ++j) {
            const R v1 = ldg(a->win + CGFD_HRFIELD(j) * fs0, pbase(p) + rel_1);
            if (hrkeep(j)) lds[slot(p)][CGFD_HRFIELD(j)][hrow][O + tx] = (LT)(hv1 ? v1 : (R)0);
        }
#pragma unroll
        for (int j = 0; j < NH; ++j) {
            const R v2 = ldg(a->win + hfield(j) * fs0, pbase(p) + rel_2);
            if (hkeep(j)) lds[slot(p)][hfield(j)][O + hr2][hcol] = (LT)(hv2 ? v2 : (R)0);
        }
    }

    R lead[NF], h1[NR], h2[NH];
#pragma unroll
    for (int c = 0; c < NF; ++c) {
        lead[c] = ldg(a->win + c * fs0, pbase(z0 + D) + rel_c);
        if (c < NR) h1[c] = ldg(a->win + CGFD_HRFIELD(c) * fs0, pbase(z0 + HD) + rel_1);
    }
#pragma unroll
    for (int j = 0; j < NH; ++j) h2[j] = ldg(a->win + hfield(j) * fs0, pbase(z0 + HD) + rel_2);

    for (int z = z0; z < z1; ++z) {
        // The field stride is laundered through an empty asm once per plane,
        // so the ~40 per-field base addresses derived from it are rebuilt on
        // the scalar unit each plane instead of being held (and spilled)
        // across the whole sweep.
        long fs = g.fs;
        asm volatile("" : "+s"(fs));
        AP al = a0;
        asm volatile("" : "+s"(al));
        // Every lane loads its cell inputs and stores its results without a
        // branch: a lane outside the grid addresses its own zero guard cell
        // (x in [ni, nxp) or a clamped guard row) and reads zeros; its stores
        // are dropped (rk_store).  With no skippable memory operation between
        // them, the compiler's wait counts let this plane's cell loads stay in
        // flight while the previous plane's prefetch is consumed.
        const unsigned pbz = pbase(z) + rel_c;
        CellIn<R> cin;
        load_cell<R, ST, (HM & 1)>(al, pbz, fs, cin);
        __syncthreads(); // the previous plane's LDS reads are done (its tile is slot(z+HD) now)
        const int sc = slot(z), sh = slot(z + HD);
#pragma unroll
        for (int c = 0; c < NF; ++c) {
            lds[sc][c][O + ty][O + tx] = (LT)w[c][CI];
            if (c < NR && hrkeep(c)) lds[sh][CGFD_HRFIELD(c)][hrow][O + tx] = (LT)(hv1 ? h1[c] : (R)0);
            w[c][4] = lead[c];
        }
#pragma unroll
        for (int j = 0; j < NH; ++j)
            if (hkeep(j)) lds[sh][hfield(j)][O + hr2][hcol] = (LT)(hv2 ? h2[j] : (R)0);
        __syncthreads();
        // next plane's lead (plane z+1+D <= nk+3) and halo (plane z+1+HD <=
        // nk+3) lie inside the 4 zero planes above the grid
#pragma unroll
        for (int c = 0; c < NF; ++c) {
            lead[c] = ldg(al->win + c * fs, pbase(z + 1 + D) + rel_c);
            if (c < NR) h1[c] = ldg(al->win + CGFD_HRFIELD(c) * fs, pbase(z + 1 + HD) + rel_1);
        }
#pragma unroll
        for (int j = 0; j < NH; ++j) h2[j] = ldg(al->win + hfield(j) * fs, pbase(z + 1 + HD) + rel_2);
        R hv[NF];
#pragma unroll
        for (int c = 0; c < NF; ++c) hv[c] = (R)0;
        if (active) {
            R Dx[NF], Dy[NF], Dz[NF];
#pragma unroll
            for (int c = 0; c < NF; ++c) {
                const LT ctr = (LT)w[c][CI];
                const LT *rx = &lds[sc][c][O + ty][tx]; // taps x-O .. x-O+4
                const LT *ry = &lds[sc][c][ty][O + tx]; // taps y-O .. y-O+4
                if (SX > 0) {
                    Dx[c] = dfd<R, 1>(rx[0], ctr, rx[2], rx[3], rx[4], h, rh);
                    Dy[c] = dfd<R, 1>(ry[0], ctr, ry[2 * HX], ry[3 * HX], ry[4 * HX], h, rh);
                } else {
                    Dx[c] = dfd<R, -1>(rx[0], rx[1], rx[2], ctr, rx[4], h, rh);
                    Dy[c] = dfd<R, -1>(ry[0], ry[HX], ry[2 * HX], ctr, ry[4 * HX], h, rh);
                }
                Dz[c] = dfd<R, SZ>(w[c][0], w[c][1], w[c][2], w[c][3], w[c][4], h, rh);
            }
            cell_rates<R, SX, SZ, ST, (HM >> 1)>(al, k, cin, Dx, Dy, Dz, w, x, y, z, e, tindex, hv);
        }
        // lanes outside the grid: their buffer stores are dropped (FP32), or
        // their zero inputs and zero rates store zeros (FP64)
        rk_store<R, SZ, ST>(al, cin, w, hv, pbz, fs, active);
        if (active) pgv_update<R, ST>(al, k, cin, hv, z - (g.nk - 3));
#pragma unroll
        for (int c = 0; c < NF; ++c) {
            w[c][0] = w[c][1];
            w[c][1] = w[c][2];
            w[c][2] = w[c][3];
            w[c][3] = w[c][4];
        }
    }
}

#undef CGFD_HRFIELD

// ---------------------------------------------------------------------------
// Variant "split": each stage as two kernels, one per half of the system.
//   PART 0 (velocity): derivatives of the 6 stresses -> momentum rates of
//     Vx, Vy, Vz, their PML auxiliaries, the traction-image override, point
//     forces, the RK update of V and the PGV maximum;
//   PART 1 (stress): derivatives of the 3 velocities (with the free-surface
//     closure of the top rows) -> Hooke rates of the 6 stresses, their PML
//     auxiliaries (including the surface terms), moment sources and the RK
//     update of T.
// Within a stage the two halves only read the stage input, so they are
// independent.  Each half keeps roughly half the live state of the fused
// kernel, which doubles the resident waves; the price is that both read the
// nine metrics.  Arithmetic per component is exactly that of finish_cell.
// ---------------------------------------------------------------------------
#ifndef CGFD_SPLIT_WAVES_V
#define CGFD_SPLIT_WAVES_V 4
#endif
#ifndef CGFD_SPLIT_WAVES_T
#define CGFD_SPLIT_WAVES_T 4
#endif

constexpr int SPLIT_TY = 4; // rows per workgroup of the split kernels

template <int PART> struct SplitFields;
template <> struct SplitFields<0> { // stencil inputs: the six stresses
    static constexpr int n = 6;
    __device__ static constexpr int comp(int i) { return TXX + i; }
};
template <> struct SplitFields<1> { // stencil inputs: the three velocities
    static constexpr int n = 3;
    __device__ static constexpr int comp(int i) { return VX + i; }
};

template <typename R, int SX, int SZ, int ST, int PART>
__global__ void __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 ? (PART == 0 ? CGFD_SPLIT_WAVES_V : CGFD_SPLIT_WAVES_T) : 2, 8)))
stage_split_kernel(const StageArgs<R> *pa, int tinc)
{
    typedef const __attribute__((address_space(4))) StageArgs<R> *AP;
    constexpr int NIN = SplitFields<PART>::n;
    constexpr int TY = SPLIT_TY, HX = 64 + 4, HY = TY + 4;
    constexpr int O = SX > 0 ? 1 : 3;
    constexpr int ZOFF = SZ > 0 ? -1 : -3;
    constexpr int CI = SZ > 0 ? 1 : 3;
    __shared__ R lds[NIN][HY][HX];

    const AP a0 = (AP)pa;
    const int tindex = 2 * *a0->step + tinc;
    const Geom g = geom_of(a0);
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * 64 + tx;
    const int ntx = (g.ni + 63) / 64, nty = (g.nj + TY - 1) / TY;
    const int nb = gridDim.x, id = blockIdx.x;
    const int q = nb / 8, r = nb % 8, xcd = id % 8;
    const int L = xcd * q + min(xcd, r) + id / 8;
    const int bx = L % ntx, by = (L / ntx) % nty, bz = L / (ntx * nty);
    const int x0 = bx * 64, y0 = by * TY;
    const int x = x0 + tx, y = y0 + ty;
    const bool active = x < g.ni && y < g.nj;
    const int zlen = a0->zlen;
    const int z0 = bz * zlen;
    const int z1 = min(g.nk, z0 + zlen);
    if (z0 >= z1) return;

    const double h = a0->h, rh = a0->rh;
    const unsigned esz = sizeof(R);
    const unsigned plb = (unsigned)(g.plane * esz);
    const int hk = tid >> 6;
    const int hrow = SX > 0 ? (hk == 0 ? 0 : TY + hk) : (hk < 3 ? hk : TY + 3);
    const int hy = y0 - O + hrow, hxx = x0 + tx;
    const bool hv1 = hy >= -3 && hy < g.nj + 3 && hxx < g.ni + 3;
    const int hc2k = tid >> 2, hr2 = tid & 3;
    const int hcol = SX > 0 ? (hc2k == 0 ? 0 : 64 + hc2k) : (hc2k < 3 ? hc2k : 67);
    const int hx2 = x0 - O + hcol, hy2 = y0 + hr2;
    const bool hv2 = tid < 16 && hx2 >= -3 && hx2 < g.ni + 3 && hy2 < g.nj + 3;
    const int rel_c = (y * g.nxp + x) * (int)esz;
    const int rel_1 = (hy * g.nxp + hxx) * (int)esz;
    const int rel_2 = (hy2 * g.nxp + hx2) * (int)esz;
    const unsigned base0 = (unsigned)(g.base * esz);
    auto pbase = [&](int zz) { return base0 + (unsigned)zz * plb; };

    const Column<R> k = column_of<R>(a0, active ? x : 0, active ? y : 0);
    int e = active ? first_source(a0, k.col, z0) : -1;

    const long fs0 = g.fs;
    R w[NIN][5];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int i = 0; i < NIN; ++i)
            w[i][j] = ldg(a0->win + SplitFields<PART>::comp(i) * fs0, pbase(z0 + ZOFF + j) + rel_c);
    R lead[NIN], h1[NIN];
#pragma unroll
    for (int i = 0; i < NIN; ++i) {
        const R *f = a0->win + SplitFields<PART>::comp(i) * fs0;
        lead[i] = ldg(f, pbase(z0 + ZOFF + 4) + rel_c);
        h1[i] = hv1 ? ldg(f, pbase(z0) + rel_1) : (R)0;
    }

    for (int z = z0; z < z1; ++z) {
        long fs = g.fs;
        asm volatile("" : "+s"(fs));
        AP a = a0;
        asm volatile("" : "+s"(a));
        const unsigned pb = pbase(z) + rel_c;
        const int top = z - (g.nk - 3);

        // per-cell inputs, issued before the stencil work
        R m[9];
        R med0 = 0, med1 = 0;
        constexpr int NO = PART == 0 ? 3 : 6; // output components
        constexpr int C0 = PART == 0 ? VX : TXX;
        R fw[NO], tw[NO];
        if (active) {
#pragma unroll
            for (int i = 0; i < 9; ++i) m[i] = ldg(a->met + i * fs, pb);
            if (PART == 0) {
                med0 = ldg(a->med + 2 * fs, pb); // 1/rho
            } else {
                med0 = ldg(a->med, pb);      // lam2mu
                med1 = ldg(a->med + fs, pb); // miu
            }
#pragma unroll
            for (int c = 0; c < NO; ++c) {
                if (ST == 1 || ST == 2) fw[c] = ldg(a->fw + (C0 + c) * fs, pb);
                if (ST >= 1) tw[c] = ldg((const R *)a->tw + (C0 + c) * fs, pb);
            }
        }
        // the x-halo column values of this plane (16 lanes)
        R h2[NIN];
#pragma unroll
        for (int i = 0; i < NIN; ++i) h2[i] = hv2 ? ldg(a->win + SplitFields<PART>::comp(i) * fs, pbase(z) + rel_2) : (R)0;

        __syncthreads();
#pragma unroll
        for (int i = 0; i < NIN; ++i) {
            lds[i][O + ty][O + tx] = w[i][CI];
            lds[i][hrow][O + tx] = h1[i];
            if (tid < 16) lds[i][O + hr2][hcol] = h2[i];
            w[i][4] = lead[i];
        }
        __syncthreads();
        if (z + 1 < z1) {
#pragma unroll
            for (int i = 0; i < NIN; ++i) {
                const R *f = a->win + SplitFields<PART>::comp(i) * fs;
                lead[i] = ldg(f, pbase(z + 1 + ZOFF + 4) + rel_c);
                h1[i] = hv1 ? ldg(f, pbase(z + 1) + rel_1) : (R)0;
            }
        }

        if (active) {
            R Dx[NIN], Dy[NIN], Dz[NIN];
#pragma unroll
            for (int i = 0; i < NIN; ++i) {
                const R ctr = w[i][CI];
                const R *rx = &lds[i][O + ty][tx];
                const R *ry = &lds[i][ty][O + tx];
                if (SX > 0) {
                    Dx[i] = dfd<R, 1>(rx[0], ctr, rx[2], rx[3], rx[4], h, rh);
                    Dy[i] = dfd<R, 1>(ry[0], ctr, ry[2 * HX], ry[3 * HX], ry[4 * HX], h, rh);
                } else {
                    Dx[i] = dfd<R, -1>(rx[0], rx[1], rx[2], ctr, rx[4], h, rh);
                    Dy[i] = dfd<R, -1>(ry[0], ry[HX], ry[2 * HX], ctr, ry[4 * HX], h, rh);
                }
                Dz[i] = dfd<R, SZ>(w[i][0], w[i][1], w[i][2], w[i][3], w[i][4], h, rh);
            }
            const R xix = m[XIX], xiy = m[XIY], xiz = m[XIZ];
            const R etx = m[ETX], ety = m[ETY], etz = m[ETZ];
            const R ztx = m[ZTX], zty = m[ZTY], ztz = m[ZTZ];
            const R *pmz = a->pml + 3 * g.ni + 3 * g.nj;
            const R APDz = pmz[z], Bz = pmz[g.nk + z], DBz = pmz[2 * g.nk + z];
            const R Bx = k.Bx, By = k.By;
            const bool unitB = Bx == (R)1 && By == (R)1 && Bz == (R)1;
            const unsigned tcell = (unsigned)(top > 0 ? top : 0) * (unsigned)(g.nj * g.nxp) + k.cxy;
            const bool inz = z < a->nabs[4] || z >= g.nk - a->nabs[5];
            R hv[NO];
            // directional terms: [0..NO) along xi, [NO..2NO) eta, [2NO..3NO) zeta
            R Dxi[NO], Det[NO], Dzt[NO];
            R DzVx1 = 0, DzVx2 = 0, DzVy1 = 0, DzVy2 = 0, DzVz1 = 0, DzVz2 = 0;
            R lam2mu = 0, lambda = 0, miu = 0;

            if constexpr (PART == 0) {
                // CalWaveCL momentum terms (calculate.cu:4878-4890); Dx[i] is
                // the xi-derivative of component TXX+i
                const R rho = med0;
                constexpr int XX = 0, YY = 1, ZZ = 2, XY = 3, XZ = 4, YZ = 5;
                if (z < a->zcon) { // hybrid grid below the conversion depth (calculate.cu:3910-3925)
                    Dxi[0] = rho * xix * Dx[XX];
                    Det[0] = rho * ety * Dy[XY];
                    Dzt[0] = ztz * Dz[XZ] * rho;
                    Dxi[1] = rho * xix * Dx[XY];
                    Det[1] = rho * ety * Dy[YY];
                    Dzt[1] = ztz * Dz[YZ] * rho;
                    Dxi[2] = rho * xix * Dx[XZ];
                    Det[2] = rho * ety * Dy[YZ];
                    Dzt[2] = ztz * Dz[ZZ] * rho;
                } else {
                Dxi[0] = (xix * Dx[XX] + xiy * Dx[XY] + xiz * Dx[XZ]) * rho;
                Det[0] = (etx * Dy[XX] + ety * Dy[XY] + etz * Dy[XZ]) * rho;
                Dzt[0] = (ztx * Dz[XX] + zty * Dz[XY] + ztz * Dz[XZ]) * rho;
                Dxi[1] = (xix * Dx[XY] + xiy * Dx[YY] + xiz * Dx[YZ]) * rho;
                Det[1] = (etx * Dy[XY] + ety * Dy[YY] + etz * Dy[YZ]) * rho;
                Dzt[1] = (ztx * Dz[XY] + zty * Dz[YY] + ztz * Dz[YZ]) * rho;
                Dxi[2] = (xix * Dx[XZ] + xiy * Dx[YZ] + xiz * Dx[ZZ]) * rho;
                Det[2] = (etx * Dy[XZ] + ety * Dy[YZ] + etz * Dy[ZZ]) * rho;
                Dzt[2] = (ztx * Dz[XZ] + zty * Dz[YZ] + ztz * Dz[ZZ]) * rho;
                }
            } else {
                // velocity zeta-derivatives of the top three rows (CondFreeVLOW,
                // calculate.cu:4736-4769)
                R cx[9], cy[9];
                if (top == 2) {
#pragma unroll
                    for (int mm = 0; mm < 9; ++mm) {
                        cx[mm] = a->cov[mm * a->cstr + k.cxy];
                        cy[mm] = a->cov[(9 + mm) * a->cstr + k.cxy];
                    }
                    Dz[0] = cx[0] * Dx[0] + cx[1] * Dx[1] + cx[2] * Dx[2] + cy[0] * Dy[0] + cy[1] * Dy[1] + cy[2] * Dy[2];
                    Dz[1] = cx[3] * Dx[0] + cx[4] * Dx[1] + cx[5] * Dx[2] + cy[3] * Dy[0] + cy[4] * Dy[1] + cy[5] * Dy[2];
                    Dz[2] = cx[6] * Dx[0] + cx[7] * Dx[1] + cx[8] * Dx[2] + cy[6] * Dy[0] + cy[7] * Dy[1] + cy[8] * Dy[2];
                    if (k.inx || k.iny) {
                        DzVx1 = cx[0] * Dx[0] + cx[1] * Dx[1] + cx[2] * Dx[2];
                        DzVx2 = cy[0] * Dy[0] + cy[1] * Dy[1] + cy[2] * Dy[2];
                        DzVy1 = cx[3] * Dx[0] + cx[4] * Dx[1] + cx[5] * Dx[2];
                        DzVy2 = cy[3] * Dy[0] + cy[4] * Dy[1] + cy[5] * Dy[2];
                        DzVz1 = cx[6] * Dx[0] + cx[7] * Dx[1] + cx[8] * Dx[2];
                        DzVz2 = cy[6] * Dy[0] + cy[7] * Dy[1] + cy[8] * Dy[2];
                    }
                } else if (top == 1) {
#pragma unroll
                    for (int c = 0; c < 3; ++c)
                        Dz[c] = SZ > 0 ? (w[c][2] - w[c][1]) / a->hr : (w[c][3] - w[c][2]) / a->hr;
                } else if (top == 0) {
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        double s = SZ > 0 ? M24A0 * (double)w[c][1] + M24A1 * (double)w[c][2] + M24A2 * (double)w[c][3]
                                          : M24A2 * (double)w[c][1] + M24A1 * (double)w[c][2] + M24A0 * (double)w[c][3];
                        Dz[c] = (R)(SZ > 0 ? s / h : s / (-h));
                    }
                }
                // CalWaveCL Hooke terms (calculate.cu:4892-4920)
                lam2mu = med0;
                miu = med1;
                lambda = lam2mu - (R)2 * miu;
                if (z < a->zcon) { // hybrid grid below the conversion depth (calculate.cu:3927-3956)
                    Dxi[0] = lam2mu * xix * Dx[0];
                    Det[0] = lambda * ety * Dy[1];
                    Dzt[0] = lambda * ztz * Dz[2];
                    Dxi[1] = lambda * xix * Dx[0];
                    Det[1] = lam2mu * ety * Dy[1];
                    Dzt[1] = lambda * ztz * Dz[2];
                    Dxi[2] = lambda * xix * Dx[0];
                    Det[2] = lambda * ety * Dy[1];
                    Dzt[2] = lam2mu * ztz * Dz[2];
                    Dxi[3] = miu * xix * Dx[1];
                    Det[3] = miu * ety * Dy[0];
                    Dzt[3] = (R)0;
                    Dxi[4] = miu * xix * Dx[2];
                    Det[4] = (R)0;
                    Dzt[4] = ztz * Dz[0] * miu;
                    Dxi[5] = (R)0;
                    Det[5] = miu * ety * Dy[2];
                    Dzt[5] = ztz * Dz[1] * miu;
                } else {
                Dxi[0] = lam2mu * xix * Dx[0] + lambda * xiy * Dx[1] + lambda * xiz * Dx[2];
                Det[0] = lam2mu * etx * Dy[0] + lambda * ety * Dy[1] + lambda * etz * Dy[2];
                Dzt[0] = lam2mu * ztx * Dz[0] + lambda * zty * Dz[1] + lambda * ztz * Dz[2];
                Dxi[1] = lambda * xix * Dx[0] + lam2mu * xiy * Dx[1] + lambda * xiz * Dx[2];
                Det[1] = lambda * etx * Dy[0] + lam2mu * ety * Dy[1] + lambda * etz * Dy[2];
                Dzt[1] = lambda * ztx * Dz[0] + lam2mu * zty * Dz[1] + lambda * ztz * Dz[2];
                Dxi[2] = lambda * xix * Dx[0] + lambda * xiy * Dx[1] + lam2mu * xiz * Dx[2];
                Det[2] = lambda * etx * Dy[0] + lambda * ety * Dy[1] + lam2mu * etz * Dy[2];
                Dzt[2] = lambda * ztx * Dz[0] + lambda * zty * Dz[1] + lam2mu * ztz * Dz[2];
                Dxi[3] = (xiy * Dx[0] + xix * Dx[1]) * miu;
                Det[3] = (ety * Dy[0] + etx * Dy[1]) * miu;
                Dzt[3] = (zty * Dz[0] + ztx * Dz[1]) * miu;
                Dxi[4] = (xiz * Dx[0] + xix * Dx[2]) * miu;
                Det[4] = (etz * Dy[0] + etx * Dy[2]) * miu;
                Dzt[4] = (ztz * Dz[0] + ztx * Dz[2]) * miu;
                Dxi[5] = (xiz * Dx[1] + xiy * Dx[2]) * miu;
                Det[5] = (etz * Dy[1] + ety * Dy[2]) * miu;
                Dzt[5] = (ztz * Dz[1] + zty * Dz[2]) * miu;
                }
            }
#pragma unroll
            for (int c = 0; c < NO; ++c)
                hv[c] = unitB ? Dxi[c] + Det[c] + Dzt[c] : Dxi[c] / Bx + Det[c] / By + Dzt[c] / Bz;

            // ---- ADE CFS-PML (calculate.cu:4993-5107) + auxiliary RK update ----
            auto aux_update = [&](int AX, long aidx, const R *ai, const R *ha) {
                const long as = a->astr[AX];
#pragma unroll
                for (int c = 0; c < NO; ++c) {
                    const long o = (C0 + c) * as + aidx;
                    if (ST == 0) {
                        a->aout[AX][o] = ai[c] + a->alpha * ha[c];
                        a->at[AX][o] = ai[c] + a->beta * ha[c];
                    } else if (ST < 3) {
                        a->aout[AX][o] = a->af[AX][o] + a->alpha * ha[c];
                        a->at[AX][o] = a->at[AX][o] + a->beta * ha[c];
                    } else {
                        a->aout[AX][o] = a->at[AX][o] + a->beta * ha[c];
                    }
                }
            };
            if (k.inx) {
                const long aidx = ((long)z * g.nj + y) * a->npml[0] + k.ix;
                R ai[NO], ha[NO];
#pragma unroll
                for (int c = 0; c < NO; ++c) ai[c] = a->ain[0][(C0 + c) * a->astr[0] + aidx];
#pragma unroll
                for (int c = 0; c < NO; ++c) hv[c] -= ai[c] / Bx;
#pragma unroll
                for (int c = 0; c < NO; ++c) ha[c] = Dxi[c] * k.DBx - k.APDx * ai[c];
                if constexpr (PART == 1) {
                    if (top == 2) {
                        const R DBx = k.DBx;
                        ha[0] += DBx * Bx * (lam2mu * ztx * DzVx1 + lambda * zty * DzVy1 + lambda * ztz * DzVz1);
                        ha[1] += DBx * Bx * (lambda * ztx * DzVx1 + lam2mu * zty * DzVy1 + lambda * ztz * DzVz1);
                        ha[2] += DBx * Bx * (lambda * ztx * DzVx1 + lambda * zty * DzVy1 + lam2mu * ztz * DzVz1);
                        ha[3] += DBx * Bx * (zty * DzVx1 + ztx * DzVy1) * miu;
                        ha[4] += DBx * Bx * (ztz * DzVx1 + ztx * DzVz1) * miu;
                        ha[5] += DBx * Bx * (ztz * DzVy1 + zty * DzVz1) * miu;
                    }
                } else {
                    if (top >= 0) {
#pragma unroll
                        for (int c = 0; c < 3; ++c) ha[c] = a->timg[(3 + c) * a->tstr + tcell];
                    }
                }
                aux_update(0, aidx, ai, ha);
            }
            if (k.iny) {
                const long aidx = ((long)z * a->npml[1] + k.iy) * g.nxp + x;
                R ai[NO], ha[NO];
#pragma unroll
                for (int c = 0; c < NO; ++c) ai[c] = a->ain[1][(C0 + c) * a->astr[1] + aidx];
#pragma unroll
                for (int c = 0; c < NO; ++c) hv[c] -= ai[c] / By;
#pragma unroll
                for (int c = 0; c < NO; ++c) ha[c] = Det[c] * k.DBy - k.APDy * ai[c];
                if constexpr (PART == 1) {
                    if (top == 2) {
                        const R DBy = k.DBy;
                        ha[0] += DBy * By * (lam2mu * ztx * DzVx2 + lambda * zty * DzVy2 + lambda * ztz * DzVz2);
                        ha[1] += DBy * By * (lambda * ztx * DzVx2 + lam2mu * zty * DzVy2 + lambda * ztz * DzVz2);
                        ha[2] += DBy * By * (lambda * ztx * DzVx2 + lambda * zty * DzVy2 + lam2mu * ztz * DzVz2);
                        ha[3] += DBy * By * (zty * DzVx2 + ztx * DzVy2) * miu;
                        ha[4] += DBy * By * (ztz * DzVx2 + ztx * DzVz2) * miu;
                        ha[5] += DBy * By * (ztz * DzVy2 + zty * DzVz2) * miu;
                    }
                } else {
                    if (top >= 0) {
#pragma unroll
                        for (int c = 0; c < 3; ++c) ha[c] = a->timg[(6 + c) * a->tstr + tcell];
                    }
                }
                aux_update(1, aidx, ai, ha);
            }
            if (inz) {
                const int iz = z < a->nabs[4] ? z : z - (g.nk - a->nabs[5]) + a->nabs[4];
                const long aidx = ((long)iz * g.nj + y) * g.nxp + x;
                R ai[NO], ha[NO];
#pragma unroll
                for (int c = 0; c < NO; ++c) ai[c] = a->ain[2][(C0 + c) * a->astr[2] + aidx];
#pragma unroll
                for (int c = 0; c < NO; ++c) hv[c] -= ai[c] / Bz;
#pragma unroll
                for (int c = 0; c < NO; ++c) ha[c] = Dzt[c] * DBz - APDz * ai[c];
                aux_update(2, aidx, ai, ha);
            }

            if constexpr (PART == 0) {
                if (top >= 0) { // CalTIMG override
#pragma unroll
                    for (int c = 0; c < 3; ++c) hv[c] = a->timg[c * a->tstr + tcell];
                }
            }

            // ---- point sources of this half, in source order ----------------------
            if (e >= 0) {
                int4 en = a->sent[e];
                while (en.w == k.col && en.x == z) {
                    const R *cf = a->scoef + (long)e * 8;
                    if (PART == 0 && en.y == 0) {
                        const R stf = a->fstf[(long)en.z * a->nstf_f + tindex];
                        const R A = cf[0];
                        hv[0] += stf * cf[1] * A;
                        hv[1] += stf * cf[2] * A;
                        hv[2] += stf * cf[3] * A;
                    } else if (PART == 1 && en.y == 1) {
                        const R stf = a->mstf[(long)en.z * a->nstf_m + tindex];
                        R A = (R)1;
                        A = stf * A / cf[0];
#pragma unroll
                        for (int c = 0; c < 6; ++c) hv[c] -= cf[1 + c] * A;
                    } else if (PART == 1 && en.y == 2) {
                        const R *fv = a->rstf + (long)en.z * 6 * a->nstf_r + tindex;
                        const R A = cf[0];
#pragma unroll
                        for (int c = 0; c < 6; ++c) hv[c] -= fv[(long)c * a->nstf_r] * A;
                    }
                    en = a->sent[++e];
                }
            }

            // ---- RK4 update ----------------------------------------------------------
#pragma unroll
            for (int c = 0; c < NO; ++c) {
                R *wo = a->wout + (C0 + c) * fs;
                R *twp = a->tw + (C0 + c) * fs;
                if (ST == 0) {
                    // the stage input is FW itself; for the stencil components it
                    // is already in the window, otherwise load it
                    R f0;
                    if (PART == 0) f0 = ldg(a->win + (VX + c) * fs, pb);
                    else f0 = ldg(a->win + (TXX + c) * fs, pb);
                    stg(wo, pb, f0 + a->alpha * hv[c]);
                    stg(twp, pb, f0 + a->beta * hv[c]);
                } else if (ST < 3) {
                    stg(wo, pb, fw[c] + a->alpha * hv[c]);
                    stg(twp, pb, tw[c] + a->beta * hv[c]);
                } else {
                    const R wn = tw[c] + a->beta * hv[c];
                    stg(wo, pb, wn);
                    if (PART == 0 && top == 2 && a->pgv) {
                        R *pv = a->pgv + c * a->cstr + k.cxy;
                        *pv = rmax(rabs(wn), rabs(*pv));
                    }
                }
            }
        }
#pragma unroll
        for (int i = 0; i < NIN; ++i) {
            w[i][0] = w[i][1];
            w[i][1] = w[i][2];
            w[i][2] = w[i][3];
            w[i][3] = w[i][4];
        }
    }
}

// ---------------------------------------------------------------------------
// traction image, top three rows (CalTIMG, calculate.cu:5428-5666)
// ---------------------------------------------------------------------------
template <typename R, int SX, int SZ, int SPLIT, int BATCH>
__global__ void __launch_bounds__(256) timg_kernel(const TimgArgs<R> av, const TimgArgs<R> *__restrict__ pa)
{
    // One grid (BATCH = 0): the arguments come by value (kernel-argument
    // segment).  A batch reads pa instead: blockIdx.z = grid * SPLIT + part,
    // and grid picks the argument block (kCtxArgBytes apart).  part is the split of
    // the nine row equations.  Fields are read one by one through the
    // constant address space, so they come in by scalar loads and the field
    // pointers keep their global address space.
    typedef const __attribute__((address_space(4))) TimgArgs<R> *TP;
    const int ctx = blockIdx.z / SPLIT, bz = blockIdx.z - ctx * SPLIT;
    const TP a = BATCH ? (TP)((const char *)pa + (long)ctx * kCtxArgBytes) : (TP)&av;
    const Geom g = geom_of(a);
    const int x = blockIdx.x * 64 + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= g.ni || y >= g.nj) return;
    const long fs = g.fs, row = g.nxp, pl = g.plane;
    const R *pmx = a->pml;
    const R *pmy = a->pml + 3 * g.ni;
    const R APDx = pmx[x], Bx = pmx[g.ni + x], DBx = pmx[2 * g.ni + x];
    const R APDy = pmy[y], By = pmy[g.nj + y], DBy = pmy[2 * g.nj + y];
    const bool inx = x < a->nabs[0] || x >= g.ni - a->nabs[1];
    const int ix = x < a->nabs[0] ? x : x - (g.ni - a->nabs[1]) + a->nabs[0];
    const bool iny = y < a->nabs[2] || y >= g.nj - a->nabs[3];
    const int iy = y < a->nabs[2] ? y : y - (g.nj - a->nabs[3]) + a->nabs[2];
    const double hx = SX > 0 ? a->h : -a->h;
    const double hz = SZ > 0 ? a->h : -a->h;
    // momentum rows: (Txx,Txy,Txz) -> Vx, (Txy,Tyy,Tyz) -> Vy, (Txz,Tyz,Tzz) -> Vz
    constexpr int trow[3][3] = {{TXX, TXY, TXZ}, {TXY, TYY, TYZ}, {TXZ, TYZ, TZZ}};
    const R *jac = a->met + JAC * fs;

    // The 3 top rows x 3 momentum equations are independent.  Small grids
    // split them over blockIdx.z (SPLIT = 9: one pair per block; 3: one row
    // per block) so a thread does not run nine dependent load chains; large
    // grids keep all nine per thread (SPLIT = 1).  With SPLIT 1 or 3 a point's
    // metrics and six stresses are loaded once for the three momentum rows,
    // and the zeta traction vectors of the planes below the surface once for
    // all rows (the products are the same, so are the results).  The loops
    // stay unrolled, so r and m are compile-time constants in every copy.
    const int rz = SPLIT == 9 ? bz / 3 : bz, mz = bz % 3;
    constexpr int TS[6] = {TXX, TYY, TZZ, TXY, TXZ, TYZ};
    // row m of the traction in terms of TS: (Txx,Txy,Txz), (Txy,Tyy,Tyz), (Txz,Tyz,Tzz)
    constexpr int tsel[3][3] = {{0, 3, 4}, {3, 1, 5}, {4, 5, 2}};
    // J * (d . T_m) for the 3 rows m at point q, d the metric triple at slot d0
    auto traction = [&](long q, int d0, R (&out)[3]) {
        R t[6];
#pragma unroll
        for (int u = 0; u < 6; ++u) t[u] = a->win[TS[u] * fs + q];
        const R j = jac[q], m0 = a->met[d0 * fs + q], m1 = a->met[(d0 + 1) * fs + q], m2 = a->met[(d0 + 2) * fs + q];
#pragma unroll
        for (int m = 0; m < 3; ++m) out[m] = j * (m0 * t[tsel[m][0]] + m1 * t[tsel[m][1]] + m2 * t[tsel[m][2]]);
    };
    // zeta traction vectors of the planes nk-6+r0 .. nk-2 (the surface plane's
    // own traction is zero and never read)
    R tz[5][3];
    const int r0 = SPLIT == 3 ? rz : 0;
    if (SPLIT != 9) {
#pragma unroll
        for (int j = 0; j < 5; ++j)
            if (j < 5 - r0) traction(gidx(g, x, y, g.nk - 6 + r0 + j), ZTX, tz[j]);
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        if (SPLIT > 1 && r != rz) continue;
        const int z = g.nk - 3 + r;
        const long G = gidx(g, x, y, z);
        const R rhojac = (R)(1.0 / (double)a->med[3 * fs + G] / (double)jac[G]);
        const long tcell = ((long)r * g.nj + y) * row + x;
        const int s = 3 + (2 - r); // the surface in the z stencil
        R vx[3][7], vy[3][7], vz[3][7];
        if (SPLIT != 9) {
#pragma unroll
            for (int i = 0; i < 7; ++i) {
                R ox[3] = {0, 0, 0}, oy[3] = {0, 0, 0};
                if ((SX > 0 && i >= 2) || (SX < 0 && i <= 4)) {
                    traction(G + (i - 3), XIX, ox);
                    traction(G + (i - 3) * row, ETX, oy);
                }
#pragma unroll
                for (int m = 0; m < 3; ++m) {
                    vx[m][i] = ox[m];
                    vy[m][i] = oy[m];
                    // plane z-3+i is tz slot i (SPLIT 3: r0 = r) or r + i
                    vz[m][i] = i < s ? tz[SPLIT == 3 ? i : r + i][m] : (R)0;
                }
            }
        }
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            if (SPLIT == 9 && m != mz) continue;
            if (SPLIT == 9) {
                const R *T0 = a->win + trow[m][0] * fs;
                const R *T1 = a->win + trow[m][1] * fs;
                const R *T2 = a->win + trow[m][2] * fs;
#pragma unroll
                for (int i = 0; i < 7; ++i) {
                    vx[m][i] = vy[m][i] = 0;
                    if ((SX > 0 && i >= 2) || (SX < 0 && i <= 4)) {
                        long q = G + (i - 3);
                        vx[m][i] = jac[q] * (a->met[XIX * fs + q] * T0[q] + a->met[XIY * fs + q] * T1[q] +
                                             a->met[XIZ * fs + q] * T2[q]);
                        q = G + (i - 3) * row;
                        vy[m][i] = jac[q] * (a->met[ETX * fs + q] * T0[q] + a->met[ETY * fs + q] * T1[q] +
                                             a->met[ETZ * fs + q] * T2[q]);
                    }
                    vz[m][i] = 0;
                    if (i < s) {
                        const long q = G + (i - 3) * pl;
                        vz[m][i] = jac[q] * (a->met[ZTX * fs + q] * T0[q] + a->met[ZTY * fs + q] * T1[q] +
                                             a->met[ZTZ * fs + q] * T2[q]);
                    }
                }
            }
            // odd images above the surface, zero traction on it
#pragma unroll
            for (int j = 1; j <= 3; ++j)
                if (j <= 3 - (2 - r)) vz[m][s + j] = (R)0 - vz[m][s - j];
            vz[m][s] = 0;

            const R DxTx = (R)((double)rhojac * (drp_sum_strf<R, SX>(vx[m]) / hx));
            const R DyTy = (R)((double)rhojac * (drp_sum_strf<R, SX>(vy[m]) / hx));
            const R DzTz = (R)((double)rhojac * (drp_sum_strf<R, SZ>(vz[m]) / hz));
            R hv = (Bx == (R)1 && By == (R)1) ? DxTx + DyTy + DzTz : DxTx / Bx + DyTy / By + DzTz;
            if (inx) {
                const R av = a->ain[0][m * a->astr[0] + ((long)z * g.nj + y) * a->npml[0] + ix];
                hv -= av / Bx;
                a->timg[(3 + m) * a->tstr + tcell] = DBx * DxTx - APDx * av;
            }
            if (iny) {
                const R av = a->ain[1][m * a->astr[1] + ((long)z * a->npml[1] + iy) * row + x];
                hv -= av / By;
                a->timg[(6 + m) * a->tstr + tcell] = DBy * DyTy - APDy * av;
            }
            a->timg[m * a->tstr + tcell] = hv;
        }
    }
}

// ---------------------------------------------------------------------------
// free-surface velocity closure matrices (VelPDcoeff, calculate.cu:3320-3466)
// ---------------------------------------------------------------------------
template <typename R> __device__ void matinv3(R *A)
{
    for (int i = 0; i < 3; i++) {
        R con = A[i * 3 + i];
        A[i * 3 + i] = 1;
        for (int j = 0; j < 3; j++) A[i * 3 + j] = A[i * 3 + j] / con;
        for (int j = 0; j < 3; j++)
            if (j != i) {
                con = A[j * 3 + i];
                A[j * 3 + i] = 0;
                for (int k = 0; k < 3; k++) A[j * 3 + k] = A[j * 3 + k] - A[i * 3 + k] * con;
            }
    }
}

template <typename R> __device__ void matmul3(const R *A, const R *B, R *C)
{
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            C[i * 3 + j] = (R)0.0;
            for (int k = 0; k < 3; k++) C[i * 3 + j] = C[i * 3 + j] + A[i * 3 + k] * B[k * 3 + j];
        }
}

template <typename R>
__global__ void velpd_kernel(Geom g, const R *met, const R *med, const R *pml, R *cov, long cstr)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= g.ni || y >= g.nj) return;
    const int z = g.nk - 1;
    const long G = gidx(g, x, y, z);
    const long fs = g.fs;
    const R *pmx = pml, *pmy = pml + 3 * g.ni, *pmz = pmy + 3 * g.nj;
    const R Bzx = pmz[g.nk + z] / pmx[g.ni + x];
    const R Bzy = pmz[g.nk + z] / pmy[g.nj + y];
    const R miu = med[fs + G];    // rho*beta*beta
    const R lam2mu = med[G];      // rho*alpha*alpha
    const R lambda = lam2mu - (R)2 * miu;
    const R e11 = met[XIX * fs + G], e12 = met[XIY * fs + G], e13 = met[XIZ * fs + G];
    const R e21 = met[ETX * fs + G], e22 = met[ETY * fs + G], e23 = met[ETZ * fs + G];
    const R e31 = met[ZTX * fs + G], e32 = met[ZTY * fs + G], e33 = met[ZTZ * fs + G];
    R A[9], B[9], C[9], t[9];
    A[0] = lam2mu * e31 * e31 + miu * (e32 * e32 + e33 * e33);
    A[1] = lambda * e31 * e32 + miu * e32 * e31;
    A[2] = lambda * e31 * e33 + miu * e33 * e31;
    A[3] = lambda * e32 * e31 + miu * e31 * e32;
    A[4] = lam2mu * e32 * e32 + miu * (e31 * e31 + e33 * e33);
    A[5] = lambda * e32 * e33 + miu * e33 * e32;
    A[6] = lambda * e33 * e31 + miu * e31 * e33;
    A[7] = lambda * e33 * e32 + miu * e32 * e33;
    A[8] = lam2mu * e33 * e33 + miu * (e31 * e31 + e32 * e32);
    matinv3(A);
    B[0] = lam2mu * e31 * e11 + miu * (e32 * e12 + e33 * e13);
    B[1] = lambda * e31 * e12 + miu * e32 * e11;
    B[2] = lambda * e31 * e13 + miu * e33 * e11;
    B[3] = lambda * e32 * e11 + miu * e31 * e12;
    B[4] = lam2mu * e32 * e12 + miu * (e31 * e11 + e33 * e13);
    B[5] = lambda * e32 * e13 + miu * e33 * e12;
    B[6] = lambda * e33 * e11 + miu * e31 * e13;
    B[7] = lambda * e33 * e12 + miu * e32 * e13;
    B[8] = lam2mu * e33 * e13 + miu * (e32 * e12 + e31 * e11);
    C[0] = lam2mu * e31 * e21 + miu * (e32 * e22 + e33 * e23);
    C[1] = lambda * e31 * e22 + miu * e32 * e21;
    C[2] = lambda * e31 * e23 + miu * e33 * e21;
    C[3] = lambda * e32 * e21 + miu * e31 * e22;
    C[4] = lam2mu * e32 * e22 + miu * (e31 * e21 + e33 * e23);
    C[5] = lambda * e32 * e23 + miu * e33 * e22;
    C[6] = lambda * e33 * e21 + miu * e31 * e23;
    C[7] = lambda * e33 * e22 + miu * e32 * e23;
    C[8] = lam2mu * e33 * e23 + miu * (e31 * e21 + e32 * e22);
    const long cxy = (long)y * g.nxp + x;
    matmul3(A, B, t);
    for (int i = 0; i < 9; i++) cov[i * cstr + cxy] = -t[i] * Bzx;
    matmul3(A, C, t);
    for (int i = 0; i < 9; i++) cov[(9 + i) * cstr + cxy] = -t[i] * Bzy;
}

// ---------------------------------------------------------------------------
// layout conversion: reference [x][y][z] (+3 ghosts) <-> padded [z][y][x]
// ---------------------------------------------------------------------------
template <typename R> __global__ void to_padded_kernel(const R *src, R *dst, Geom g)
{
    __shared__ R tile[32][33];
    const int NY = g.nj + 6, NZ = g.nk + 6;
    const int z0 = blockIdx.x * 32, x0 = blockIdx.y * 32, y = blockIdx.z;
    const int tx = threadIdx.x, ty = threadIdx.y;
    for (int i = ty; i < 32; i += 8) {
        const int xs = x0 + i, zs = z0 + tx;
        if (xs < g.ni && zs < g.nk) tile[i][tx] = src[((long)(xs + 3) * NY + (y + 3)) * NZ + zs + 3];
    }
    __syncthreads();
    for (int i = ty; i < 32; i += 8) {
        const int zd = z0 + i, xd = x0 + tx;
        if (xd < g.ni && zd < g.nk) dst[gidx(g, xd, y, zd)] = tile[tx][i];
    }
}

template <typename R> __global__ void from_padded_kernel(const R *src, R *dst, Geom g)
{
    __shared__ R tile[32][33];
    const int NY = g.nj + 6, NZ = g.nk + 6;
    const int z0 = blockIdx.x * 32, x0 = blockIdx.y * 32, y = blockIdx.z;
    const int tx = threadIdx.x, ty = threadIdx.y;
    for (int i = ty; i < 32; i += 8) {
        const int zs = z0 + i, xs = x0 + tx;
        if (xs < g.ni && zs < g.nk) tile[i][tx] = src[gidx(g, xs, y, zs)];
    }
    __syncthreads();
    for (int i = ty; i < 32; i += 8) {
        const int xd = x0 + i, zd = z0 + tx;
        if (xd < g.ni && zd < g.nk) dst[((long)(xd + 3) * NY + (y + 3)) * NZ + zd + 3] = tile[tx][i];
    }
}

// ---------------------------------------------------------------------------
// media: alpha, beta, rho -> lam2mu, miu, 1/rho (slots 0-2), rho kept in slot 3
// ---------------------------------------------------------------------------
template <typename R> __global__ void media_prep_kernel(Geom g, R *med)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y, z = blockIdx.z;
    if (x >= g.ni) return;
    const long G = gidx(g, x, y, z);
    const R alpha = med[G], beta = med[g.fs + G], rho = med[3 * g.fs + G];
    // calculate.cu:4855-4858: miu = rho*beta*beta, lam2mu = rho*alpha*alpha,
    // and 1.0/rho (a double quotient of floats rounded to float equals the
    // float quotient)
    med[G] = rho * alpha * alpha;
    med[g.fs + G] = rho * beta * beta;
    med[2 * g.fs + G] = (R)1 / rho;
}

template <typename R> void launch_media_prep(const Geom &g, R *med, hipStream_t s)
{
    media_prep_kernel<R><<<dim3((g.ni + 255) / 256, g.nj, g.nk), dim3(256), 0, s>>>(g, med);
}

// ---------------------------------------------------------------------------
// source coefficients (the step-independent parts of LoadForce/LoadMoment)
// ---------------------------------------------------------------------------
template <typename R>
__global__ void src_prep_kernel(Geom g, const R *met, const R *med, const int4 *sent, const R *fvec,
                                const R *mvec, R *scoef, int nent, R hr)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nent) return;
    const int4 en = sent[e];
    const int x = en.w % g.ni, y = en.w / g.ni, z = en.x;
    const long G = gidx(g, x, y, z);
    const R jac = met[JAC * g.fs + G];
    R *cf = scoef + (long)e * 8;
    if (en.y == 0) {
        // calculate.cu:6173-6180
        R A = hr * hr * hr * jac;
        const R d = (R)1.0;
        A = (R)1 / (A * med[3 * g.fs + G]) * d;
        if (z == g.nk - 1) A = A * (R)2;
        cf[0] = A;
        for (int k = 0; k < 3; ++k) cf[1 + k] = fvec[3 * en.z + k];
    } else if (en.y == 1) {
        // calculate.cu:6238-6240: the divisor h*h*h*jac
        cf[0] = hr * hr * hr * jac;
        for (int k = 0; k < 6; ++k) cf[1 + k] = mvec[6 * en.z + k];
    } else {
        // LoadRmom, calculate.cu:6314-6316: A = 1/(h*h*h*jac)
        R A = (R)1;
        A = A / (hr * hr * hr * jac);
        cf[0] = A;
    }
}

// receivers (PWpick, calculate.cu:2780-2828): all 9 components of FW
template <typename R>
__global__ void record_kernel(const R *fw, long fs, const long *rlin, int nrecv, R *seis, int nt, const int *step)
{
    const int t = *step;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nrecv * NF) return;
    const int r = i / NF, c = i % NF;
    seis[((long)c * nrecv + r) * nt + t] = fw[c * fs + rlin[r]];
}

// snapshots (SWpick, calculate.cu:2360-2444): after step t with t % tinv == 0,
// ncomp components from c0 of FW at npts points into time slot t/tinv, laid
// out [comp][point]; other steps return at once
template <typename R>
__global__ void snap_kernel(const R *fw, long fs, const long *lin, int npts, int c0, int ncomp, R *buf,
                            const int *step, int tinv, int ntime)
{
    const int t = *step;
    if (t % tinv != 0 || t / tinv >= ntime) return;
    R *slot = buf + (long)(t / tinv) * ncomp * npts;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)npts * ncomp) return;
    const int c = (int)(i / npts), p = (int)(i % npts);
    slot[i] = fw[(c0 + c) * fs + lin[p]];
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
// The LDS kernel at the tile width a.tile (32 or 64 columns; FP64 always 64).
template <typename R, int SX, int SZ, int ST, int HM>
static void launch_lds(const StageArgs<R> &a, const StageArgs<R> *pa, const int *starts, int nctx, int nblocks,
                       int tinc, int zchunks, hipStream_t s)
{
    constexpr int ty = lds_ty<R>();
    const dim3 block(64, ty);
    if constexpr (sizeof(R) == 4) {
        if (a.tile == 32) {
            const int ntx = (a.g.ni + 31) / 32, nty = (a.g.nj + 2 * ty - 1) / (2 * ty);
            const dim3 grid((unsigned)(nblocks > 0 ? nblocks : ntx * nty * zchunks));
            stage_lds_kernel<R, SX, SZ, ST, HM, 32><<<grid, block, 0, s>>>(pa, starts, nctx, tinc);
            return;
        }
    }
    const int ntx = (a.g.ni + 63) / 64, nty = (a.g.nj + ty - 1) / ty;
    const dim3 grid((unsigned)(nblocks > 0 ? nblocks : ntx * nty * zchunks));
    stage_lds_kernel<R, SX, SZ, ST, HM, 64><<<grid, block, 0, s>>>(pa, starts, nctx, tinc);
}

template <typename R, int SX, int SZ, int ST>
static void launch_stage_one(int variant, const StageArgs<R> &a, const StageArgs<R> *pa, const int *starts, int nctx,
                             int nblocks, int tinc, int rows, int zchunks, hipStream_t s)
{
    if (variant == 0) {
        // mode bit 0: constant horizontal metrics; bit 1: a hybrid grid's
        // Cartesian planes
        switch ((a.hmet ? 1 : 0) | (a.zcon > 0 ? 2 : 0)) {
        case 0: launch_lds<R, SX, SZ, ST, 0>(a, pa, starts, nctx, nblocks, tinc, zchunks, s); break;
        case 1: launch_lds<R, SX, SZ, ST, 1>(a, pa, starts, nctx, nblocks, tinc, zchunks, s); break;
        case 2: launch_lds<R, SX, SZ, ST, 2>(a, pa, starts, nctx, nblocks, tinc, zchunks, s); break;
        default: launch_lds<R, SX, SZ, ST, 3>(a, pa, starts, nctx, nblocks, tinc, zchunks, s); break;
        }
    } else if (variant == 2) {
        const int ntx = (a.g.ni + 63) / 64, nty = (a.g.nj + SPLIT_TY - 1) / SPLIT_TY;
        stage_split_kernel<R, SX, SZ, ST, 0><<<dim3(ntx * nty * zchunks), dim3(64, SPLIT_TY), 0, s>>>(pa, tinc);
        stage_split_kernel<R, SX, SZ, ST, 1><<<dim3(ntx * nty * zchunks), dim3(64, SPLIT_TY), 0, s>>>(pa, tinc);
    } else {
        dim3 grid((a.g.ni + 63) / 64, (a.g.nj + rows - 1) / rows, zchunks);
        stage_direct_kernel<R, SX, SZ, ST><<<grid, dim3(64, rows), 0, s>>>(pa, tinc);
    }
}

template <typename R, int SX, int SZ>
static void launch_stage_sxz(int variant, int st, const StageArgs<R> &a, const StageArgs<R> *pa, const int *starts,
                             int nctx, int nblocks, int tinc, int rows, int zchunks, hipStream_t s)
{
    switch (st) {
    case 0: launch_stage_one<R, SX, SZ, 0>(variant, a, pa, starts, nctx, nblocks, tinc, rows, zchunks, s); break;
    case 1: launch_stage_one<R, SX, SZ, 1>(variant, a, pa, starts, nctx, nblocks, tinc, rows, zchunks, s); break;
    case 2: launch_stage_one<R, SX, SZ, 2>(variant, a, pa, starts, nctx, nblocks, tinc, rows, zchunks, s); break;
    default: launch_stage_one<R, SX, SZ, 3>(variant, a, pa, starts, nctx, nblocks, tinc, rows, zchunks, s); break;
    }
}

template <typename R>
void launch_stage(int variant, int sx, int sz, int st, const StageArgs<R> &a, const StageArgs<R> *pa, int tinc,
                  int rows, int zchunks, hipStream_t s)
{
    if (sx > 0 && sz > 0) launch_stage_sxz<R, 1, 1>(variant, st, a, pa, nullptr, 1, 0, tinc, rows, zchunks, s);
    else if (sx > 0) launch_stage_sxz<R, 1, -1>(variant, st, a, pa, nullptr, 1, 0, tinc, rows, zchunks, s);
    else if (sz > 0) launch_stage_sxz<R, -1, 1>(variant, st, a, pa, nullptr, 1, 0, tinc, rows, zchunks, s);
    else launch_stage_sxz<R, -1, -1>(variant, st, a, pa, nullptr, 1, 0, tinc, rows, zchunks, s);
}

template <typename R>
void launch_stage_batch(int sx, int sz, int st, int hmet, int hybrid, int tw, const StageArgs<R> *pa,
                        const int *starts, int nctx, int nblocks, int tinc, hipStream_t s)
{
    if (nctx == 0 || nctx == -1 || nblocks < 1) return;
    StageArgs<R> a{}; // only the metric mode, the hybrid flag and the tile width steer a batched launch
    a.hmet = hmet;
    a.zcon = hybrid ? 1 : 0;
    a.tile = tw;
    if (sx > 0 && sz > 0) launch_stage_sxz<R, 1, 1>(0, st, a, pa, starts, nctx, nblocks, tinc, 0, 0, s);
    else if (sx > 0) launch_stage_sxz<R, 1, -1>(0, st, a, pa, starts, nctx, nblocks, tinc, 0, 0, s);
    else if (sz > 0) launch_stage_sxz<R, -1, 1>(0, st, a, pa, starts, nctx, nblocks, tinc, 0, 0, s);
    else launch_stage_sxz<R, -1, -1>(0, st, a, pa, starts, nctx, nblocks, tinc, 0, 0, s);
}

template <typename R, int SPLIT>
static void launch_timg_split(int sx, int sz, const TimgArgs<R> &av, const TimgArgs<R> *pa, dim3 grid, dim3 block,
                              hipStream_t s)
{
    if (pa) {
        if (sx > 0 && sz > 0) timg_kernel<R, 1, 1, SPLIT, 1><<<grid, block, 0, s>>>(av, pa);
        else if (sx > 0) timg_kernel<R, 1, -1, SPLIT, 1><<<grid, block, 0, s>>>(av, pa);
        else if (sz > 0) timg_kernel<R, -1, 1, SPLIT, 1><<<grid, block, 0, s>>>(av, pa);
        else timg_kernel<R, -1, -1, SPLIT, 1><<<grid, block, 0, s>>>(av, pa);
    } else {
        if (sx > 0 && sz > 0) timg_kernel<R, 1, 1, SPLIT, 0><<<grid, block, 0, s>>>(av, pa);
        else if (sx > 0) timg_kernel<R, 1, -1, SPLIT, 0><<<grid, block, 0, s>>>(av, pa);
        else if (sz > 0) timg_kernel<R, -1, 1, SPLIT, 0><<<grid, block, 0, s>>>(av, pa);
        else timg_kernel<R, -1, -1, SPLIT, 0><<<grid, block, 0, s>>>(av, pa);
    }
}

template <typename R>
void launch_timg(int sx, int sz, int ni, int nj, long cols, const TimgArgs<R> &av, const TimgArgs<R> *pa, int nctx,
                 int rows, hipStream_t s)
{
    dim3 block(64, rows);
// Columns below which the 9 row equations get a block each, and below which
// each of the 3 rows gets one (measured per stage: 64^3 8.6 us at 9 parts,
// 10.8 at 3; 200x200 15.1 us at 3, 16.3 at 9; 600x600 52 us at 1 or 3).
#ifndef CGFD_TIMG_SPLIT9
#define CGFD_TIMG_SPLIT9 16384
#endif
#ifndef CGFD_TIMG_SPLIT3
#define CGFD_TIMG_SPLIT3 131072
#endif
    const int split = cols < CGFD_TIMG_SPLIT9 ? 9 : (cols < CGFD_TIMG_SPLIT3 ? 3 : 1); // see timg_kernel
    dim3 grid((ni + 63) / 64, (nj + rows - 1) / rows, split * nctx);
    if (split == 9) launch_timg_split<R, 9>(sx, sz, av, pa, grid, block, s);
    else if (split == 3) launch_timg_split<R, 3>(sx, sz, av, pa, grid, block, s);
    else launch_timg_split<R, 1>(sx, sz, av, pa, grid, block, s);
}

template <typename R>
void launch_velpd(const Geom &g, const R *met, const R *med, const R *pml, R *cov, long cstr, hipStream_t s)
{
    dim3 block(64);
    dim3 grid((g.ni + 63) / 64, g.nj);
    velpd_kernel<R><<<grid, block, 0, s>>>(g, met, med, pml, cov, cstr);
}

template <typename R> void launch_to_padded(const R *src, R *dst, const Geom &g, hipStream_t s)
{
    dim3 block(32, 8);
    dim3 grid((g.nk + 31) / 32, (g.ni + 31) / 32, g.nj);
    to_padded_kernel<R><<<grid, block, 0, s>>>(src, dst, g);
}

template <typename R> void launch_from_padded(const R *src, R *dst, const Geom &g, hipStream_t s)
{
    dim3 block(32, 8);
    dim3 grid((g.nk + 31) / 32, (g.ni + 31) / 32, g.nj);
    from_padded_kernel<R><<<grid, block, 0, s>>>(src, dst, g);
}

template <typename R>
void launch_src_prep(const Geom &g, const R *met, const R *med, const int4 *sent, const R *fvec, const R *mvec,
                     R *scoef, int nent, R hr, hipStream_t s)
{
    if (nent <= 0) return;
    src_prep_kernel<R><<<(nent + 63) / 64, 64, 0, s>>>(g, met, med, sent, fvec, mvec, scoef, nent, hr);
}

template <typename R>
void launch_snap(const R *fw, long fs, const long *lin, int npts, int c0, int ncomp, R *buf, const int *step,
                 int tinv, int ntime, hipStream_t s)
{
    const long n = (long)npts * ncomp;
    if (n <= 0) return;
    snap_kernel<R><<<(unsigned)((n + 255) / 256), 256, 0, s>>>(fw, fs, lin, npts, c0, ncomp, buf, step, tinv,
                                                                ntime);
}

template <typename R>
void launch_record(const R *fw, long fs, const long *rlin, int nrecv, R *seis, int nt, const int *step,
                   hipStream_t s)
{
    if (nrecv <= 0) return;
    const int n = nrecv * NF;
    record_kernel<R><<<(n + 255) / 256, 256, 0, s>>>(fw, fs, rlin, nrecv, seis, nt, step);
}

__global__ void step_advance_kernel(int *step)
{
    if (threadIdx.x == 0) step[0] += 1;
}

void launch_step_advance(int *step, hipStream_t s) { step_advance_kernel<<<1, 64, 0, s>>>(step); }

int stage_tile_rows(bool fp64, int tw) { return fp64 ? lds_ty<double>() : lds_ty<float>() * (tw == 32 ? 2 : 1); }

// batched trace gathers: blockIdx.y = grid, whose RecArgs block is
// kCtxArgBytes from the previous grid's
template <typename R> __global__ void record_batch_kernel(const RecArgs<R> *__restrict__ pa)
{
    typedef const __attribute__((address_space(4))) RecArgs<R> *RP;
    const RP a = (RP)((const char *)pa + (long)blockIdx.y * kCtxArgBytes);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a->nrecv * NF) return;
    const int t = *a->step;
    if (t >= a->nt) return;
    const int r = i / NF, c = i % NF;
    a->seis[((long)c * a->nrecv + r) * a->nt + t] = a->fw[c * a->fs + a->rlin[r]];
}

template <typename R> void launch_record_batch(const RecArgs<R> *pa, int nctx, int max_nrecv, hipStream_t s)
{
    if (max_nrecv <= 0 || nctx <= 0) return;
    const int n = max_nrecv * NF;
    record_batch_kernel<R><<<dim3((n + 255) / 256, nctx), 256, 0, s>>>(pa);
}

// every grid's step counter += 1: lane i reads grid i's counter address from
// its RecArgs block (the same offset in both precisions)
__global__ void step_advance_batch_kernel(const char *__restrict__ pa, int nctx)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nctx) return;
    int *step = ((const RecArgs<float> *)(pa + (long)i * kCtxArgBytes))->step;
    step[0] += 1;
}

void launch_step_advance_batch(const char *pa, int nctx, hipStream_t s)
{
    if (nctx <= 0) return;
    step_advance_batch_kernel<<<(nctx + 63) / 64, 64, 0, s>>>(pa, nctx);
}

#define CGFD_INSTANTIATE(R)                                                                                   \
    template void launch_stage<R>(int, int, int, int, const StageArgs<R> &, const StageArgs<R> *, int, int, int, \
                                   hipStream_t);                                                              \
    template void launch_stage_batch<R>(int, int, int, int, int, int, const StageArgs<R> *, const int *, int, int, \
                                        int, hipStream_t);                                                    \
    template void launch_timg<R>(int, int, int, int, long, const TimgArgs<R> &, const TimgArgs<R> *, int, int, \
                                  hipStream_t);                                                               \
    template void launch_record_batch<R>(const RecArgs<R> *, int, int, hipStream_t);                          \
    template void launch_velpd<R>(const Geom &, const R *, const R *, const R *, R *, long, hipStream_t);     \
    template void launch_to_padded<R>(const R *, R *, const Geom &, hipStream_t);                            \
    template void launch_from_padded<R>(const R *, R *, const Geom &, hipStream_t);                          \
    template void launch_src_prep<R>(const Geom &, const R *, const R *, const int4 *, const R *, const R *,  \
                                     R *, int, R, hipStream_t);                                               \
    template void launch_record<R>(const R *, long, const long *, int, R *, int, const int *, hipStream_t); \
    template void launch_snap<R>(const R *, long, const long *, int, int, int, R *, const int *, int, int,   \
                                 hipStream_t);                                                              \
    template void launch_media_prep<R>(const Geom &, R *, hipStream_t);

CGFD_INSTANTIATE(float)
CGFD_INSTANTIATE(double)

} // namespace cgfd
