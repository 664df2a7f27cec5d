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
template <typename R, int SX, int SZ, int ST>
__global__ void __launch_bounds__(256) stage_kernel(StageArgs<R> a)
{
    const Geom g = a.g;
    const int x = blockIdx.x * 64 + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= g.ni || y >= g.nj) return;
    const int z0 = blockIdx.z * a.zlen;
    const int z1 = min(g.nk, z0 + a.zlen);
    if (z0 >= z1) return;

    const long fs = g.fs;
    const long row = g.nxp;
    const long pl = g.plane;
    const double h = a.h, rh = a.rh;
    constexpr int ZOFF = SZ > 0 ? -1 : -3; // window slot 0 sits at z+ZOFF
    constexpr int CI = SZ > 0 ? 1 : 3;     // window slot of the cell itself

    // ---- per-column constants --------------------------------------------
    const R *pmx = a.pml;
    const R *pmy = a.pml + 3 * g.ni;
    const R *pmz = pmy + 3 * g.nj;
    const R APDx = pmx[x], Bx = pmx[g.ni + x], DBx = pmx[2 * g.ni + x];
    const R APDy = pmy[y], By = pmy[g.nj + y], DBy = pmy[2 * g.nj + y];
    const bool inx = x < a.nabs[0] || x >= g.ni - a.nabs[1];
    const int ix = x < a.nabs[0] ? x : x - (g.ni - a.nabs[1]) + a.nabs[0];
    const bool iny = y < a.nabs[2] || y >= g.nj - a.nabs[3];
    const int iy = y < a.nabs[2] ? y : y - (g.nj - a.nabs[3]) + a.nabs[2];
    const long cxy = (long)y * row + x; // (x,y) offset inside one plane-sized map

    const int col = y * g.ni + x;
    int e = a.col_first ? a.col_first[col] : -1;
    if (e >= 0)
        while (a.sent[e].w == col && a.sent[e].x < z0) ++e;

    // ---- z register window -------------------------------------------------
    long p = gidx(g, x, y, z0);
    R w[NF][5];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int c = 0; c < NF; ++c) w[c][j] = a.win[c * fs + p + (ZOFF + j) * pl];

    for (int z = z0; z < z1; ++z, p += pl) {
#pragma unroll
        for (int c = 0; c < NF; ++c) w[c][4] = a.win[c * fs + p + (ZOFF + 4) * pl];

        // ---- CalDiffCL: one-sided derivatives along xi, eta, zeta -------------
        R Dx[NF], Dy[NF], Dz[NF];
#pragma unroll
        for (int c = 0; c < NF; ++c) {
            const R *f = a.win + c * fs + p;
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

        // velocity zeta-derivatives of the top three rows (CondFreeVLOW,
        // calculate.cu:4736-4769): closure on the surface, 2-point one row
        // below, 2-4 MacCormack two rows below
        const int top = z - (g.nk - 3);
        R cx[9], cy[9];
        if (top == 2) {
#pragma unroll
            for (int k = 0; k < 9; ++k) {
                cx[k] = a.cov[k * a.cstr + cxy];
                cy[k] = a.cov[(9 + k) * a.cstr + cxy];
            }
            Dz[VX] = cx[0] * Dx[VX] + cx[1] * Dx[VY] + cx[2] * Dx[VZ] + cy[0] * Dy[VX] + cy[1] * Dy[VY] +
                     cy[2] * Dy[VZ];
            Dz[VY] = cx[3] * Dx[VX] + cx[4] * Dx[VY] + cx[5] * Dx[VZ] + cy[3] * Dy[VX] + cy[4] * Dy[VY] +
                     cy[5] * Dy[VZ];
            Dz[VZ] = cx[6] * Dx[VX] + cx[7] * Dx[VY] + cx[8] * Dx[VZ] + cy[6] * Dy[VX] + cy[7] * Dy[VY] +
                     cy[8] * Dy[VZ];
        } else if (top == 1) {
#pragma unroll
            for (int c = VX; c <= VZ; ++c)
                Dz[c] = SZ > 0 ? (w[c][2] - w[c][1]) / a.hr : (w[c][3] - w[c][2]) / a.hr;
        } else if (top == 0) {
#pragma unroll
            for (int c = VX; c <= VZ; ++c) {
                double s = SZ > 0 ? M24A0 * (double)w[c][1] + M24A1 * (double)w[c][2] + M24A2 * (double)w[c][3]
                                  : M24A2 * (double)w[c][1] + M24A1 * (double)w[c][2] + M24A0 * (double)w[c][3];
                Dz[c] = (R)(SZ > 0 ? s / h : s / (-h));
            }
        }

        // ---- CalWaveCL: contraction with metrics, Hooke, momentum -------------
        const R *mt = a.met + p;
        const R *md = a.med + p;
        R rho = md[2 * fs];
        const R miu = rho * md[fs] * md[fs];
        const R lam2mu = rho * md[0] * md[0];
        const R lambda = lam2mu - (R)2 * miu; // exact: 2*miu is exact, single rounding
        rho = (R)1 / rho;
        const R xix = mt[0], xiy = mt[fs], xiz = mt[2 * fs];
        const R etx = mt[3 * fs], ety = mt[4 * fs], etz = mt[5 * fs];
        const R ztx = mt[6 * fs], zty = mt[7 * fs], ztz = mt[8 * fs];

        const R DxiVx = (xix * Dx[TXX] + xiy * Dx[TXY] + xiz * Dx[TXZ]) * rho;
        const R DetVx = (etx * Dy[TXX] + ety * Dy[TXY] + etz * Dy[TXZ]) * rho;
        const R DztVx = (ztx * Dz[TXX] + zty * Dz[TXY] + ztz * Dz[TXZ]) * rho;
        const R DxiVy = (xix * Dx[TXY] + xiy * Dx[TYY] + xiz * Dx[TYZ]) * rho;
        const R DetVy = (etx * Dy[TXY] + ety * Dy[TYY] + etz * Dy[TYZ]) * rho;
        const R DztVy = (ztx * Dz[TXY] + zty * Dz[TYY] + ztz * Dz[TYZ]) * rho;
        const R DxiVz = (xix * Dx[TXZ] + xiy * Dx[TYZ] + xiz * Dx[TZZ]) * rho;
        const R DetVz = (etx * Dy[TXZ] + ety * Dy[TYZ] + etz * Dy[TZZ]) * rho;
        const R DztVz = (ztx * Dz[TXZ] + zty * Dz[TYZ] + ztz * Dz[TZZ]) * rho;

        const R DxiTxx = lam2mu * xix * Dx[VX] + lambda * xiy * Dx[VY] + lambda * xiz * Dx[VZ];
        const R DetTxx = lam2mu * etx * Dy[VX] + lambda * ety * Dy[VY] + lambda * etz * Dy[VZ];
        const R DztTxx = lam2mu * ztx * Dz[VX] + lambda * zty * Dz[VY] + lambda * ztz * Dz[VZ];
        const R DxiTyy = lambda * xix * Dx[VX] + lam2mu * xiy * Dx[VY] + lambda * xiz * Dx[VZ];
        const R DetTyy = lambda * etx * Dy[VX] + lam2mu * ety * Dy[VY] + lambda * etz * Dy[VZ];
        const R DztTyy = lambda * ztx * Dz[VX] + lam2mu * zty * Dz[VY] + lambda * ztz * Dz[VZ];
        const R DxiTzz = lambda * xix * Dx[VX] + lambda * xiy * Dx[VY] + lam2mu * xiz * Dx[VZ];
        const R DetTzz = lambda * etx * Dy[VX] + lambda * ety * Dy[VY] + lam2mu * etz * Dy[VZ];
        const R DztTzz = lambda * ztx * Dz[VX] + lambda * zty * Dz[VY] + lam2mu * ztz * Dz[VZ];
        const R DxiTxy = (xiy * Dx[VX] + xix * Dx[VY]) * miu;
        const R DetTxy = (ety * Dy[VX] + etx * Dy[VY]) * miu;
        const R DztTxy = (zty * Dz[VX] + ztx * Dz[VY]) * miu;
        const R DxiTxz = (xiz * Dx[VX] + xix * Dx[VZ]) * miu;
        const R DetTxz = (etz * Dy[VX] + etx * Dy[VZ]) * miu;
        const R DztTxz = (ztz * Dz[VX] + ztx * Dz[VZ]) * miu;
        const R DxiTyz = (xiz * Dx[VY] + xiy * Dx[VZ]) * miu;
        const R DetTyz = (etz * Dy[VY] + ety * Dy[VZ]) * miu;
        const R DztTyz = (ztz * Dz[VY] + zty * Dz[VZ]) * miu;

        const R APDz = pmz[z], Bz = pmz[g.nk + z], DBz = pmz[2 * g.nk + z];
        R hv[NF];
        if (Bx == (R)1 && By == (R)1 && Bz == (R)1) {
            // x/1 == x exactly: the interior skips the nine divisions
            hv[TXX] = DxiTxx + DetTxx + DztTxx;
            hv[TYY] = DxiTyy + DetTyy + DztTyy;
            hv[TZZ] = DxiTzz + DetTzz + DztTzz;
            hv[TXY] = DxiTxy + DetTxy + DztTxy;
            hv[TXZ] = DxiTxz + DetTxz + DztTxz;
            hv[TYZ] = DxiTyz + DetTyz + DztTyz;
            hv[VX] = DxiVx + DetVx + DztVx;
            hv[VY] = DxiVy + DetVy + DztVy;
            hv[VZ] = DxiVz + DetVz + DztVz;
        } else {
            hv[TXX] = DxiTxx / Bx + DetTxx / By + DztTxx / Bz;
            hv[TYY] = DxiTyy / Bx + DetTyy / By + DztTyy / Bz;
            hv[TZZ] = DxiTzz / Bx + DetTzz / By + DztTzz / Bz;
            hv[TXY] = DxiTxy / Bx + DetTxy / By + DztTxy / Bz;
            hv[TXZ] = DxiTxz / Bx + DetTxz / By + DztTxz / Bz;
            hv[TYZ] = DxiTyz / Bx + DetTyz / By + DztTyz / Bz;
            hv[VX] = DxiVx / Bx + DetVx / By + DztVx / Bz;
            hv[VY] = DxiVy / Bx + DetVy / By + DztVy / Bz;
            hv[VZ] = DxiVz / Bx + DetVz / By + DztVz / Bz;
        }

        // surface contributions routed into the x/y auxiliary stress rates
        R DzVx1 = 0, DzVx2 = 0, DzVy1 = 0, DzVy2 = 0, DzVz1 = 0, DzVz2 = 0;
        if (top == 2 && (inx || iny)) {
            DzVx1 = cx[0] * Dx[VX] + cx[1] * Dx[VY] + cx[2] * Dx[VZ];
            DzVx2 = cy[0] * Dy[VX] + cy[1] * Dy[VY] + cy[2] * Dy[VZ];
            DzVy1 = cx[3] * Dx[VX] + cx[4] * Dx[VY] + cx[5] * Dx[VZ];
            DzVy2 = cy[3] * Dy[VX] + cy[4] * Dy[VY] + cy[5] * Dy[VZ];
            DzVz1 = cx[6] * Dx[VX] + cx[7] * Dx[VY] + cx[8] * Dx[VZ];
            DzVz2 = cy[6] * Dy[VX] + cy[7] * Dy[VY] + cy[8] * Dy[VZ];
        }
        const long tcell = (long)(top > 0 ? top : 0) * g.nj * row + cxy; // top-row slot offset

        // ---- ADE CFS-PML (calculate.cu:4993-5107) + auxiliary RK update --------
        // The rates of one axis feed that axis' RK update straight away; the
        // traction image replaces the velocity rates of the top rows first.
#define CGFD_AUX_UPDATE(AX, HA, AIN)                                                                      \
    {                                                                                                     \
        const long ap = aidx;                                                                             \
        const long as = a.astr[AX];                                                                       \
        _Pragma("unroll") for (int c = 0; c < NF; ++c)                                                    \
        {                                                                                                 \
            const R hh = HA[c];                                                                           \
            if (ST == 0) {                                                                                \
                a.aout[AX][c * as + ap] = AIN[c] + a.alpha * hh;                                          \
                a.at[AX][c * as + ap] = AIN[c] + a.beta * hh;                                             \
            } else if (ST < 3) {                                                                          \
                a.aout[AX][c * as + ap] = a.af[AX][c * as + ap] + a.alpha * hh;                           \
                a.at[AX][c * as + ap] = a.at[AX][c * as + ap] + a.beta * hh;                              \
            } else {                                                                                      \
                a.aout[AX][c * as + ap] = a.at[AX][c * as + ap] + a.beta * hh;                            \
            }                                                                                             \
        }                                                                                                 \
    }

        if (inx) {
            const long aidx = ((long)z * g.nj + y) * a.npml[0] + ix;
            R ai[NF], ha[NF];
#pragma unroll
            for (int c = 0; c < NF; ++c) ai[c] = a.ain[0][c * a.astr[0] + aidx];
#pragma unroll
            for (int c = 0; c < NF; ++c) hv[c] -= ai[c] / Bx;
            ha[VX] = DxiVx * DBx - APDx * ai[VX];
            ha[VY] = DxiVy * DBx - APDx * ai[VY];
            ha[VZ] = DxiVz * DBx - APDx * ai[VZ];
            ha[TXX] = DxiTxx * DBx - APDx * ai[TXX];
            ha[TYY] = DxiTyy * DBx - APDx * ai[TYY];
            ha[TZZ] = DxiTzz * DBx - APDx * ai[TZZ];
            ha[TXY] = DxiTxy * DBx - APDx * ai[TXY];
            ha[TXZ] = DxiTxz * DBx - APDx * ai[TXZ];
            ha[TYZ] = DxiTyz * DBx - APDx * ai[TYZ];
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
                for (int c = VX; c <= VZ; ++c) ha[c] = a.timg[(3 + c) * a.tstr + tcell];
            }
            CGFD_AUX_UPDATE(0, ha, ai)
        }
        if (iny) {
            const long aidx = ((long)z * a.npml[1] + iy) * row + x;
            R ai[NF], ha[NF];
#pragma unroll
            for (int c = 0; c < NF; ++c) ai[c] = a.ain[1][c * a.astr[1] + aidx];
#pragma unroll
            for (int c = 0; c < NF; ++c) hv[c] -= ai[c] / By;
            ha[VX] = DetVx * DBy - APDy * ai[VX];
            ha[VY] = DetVy * DBy - APDy * ai[VY];
            ha[VZ] = DetVz * DBy - APDy * ai[VZ];
            ha[TXX] = DetTxx * DBy - APDy * ai[TXX];
            ha[TYY] = DetTyy * DBy - APDy * ai[TYY];
            ha[TZZ] = DetTzz * DBy - APDy * ai[TZZ];
            ha[TXY] = DetTxy * DBy - APDy * ai[TXY];
            ha[TXZ] = DetTxz * DBy - APDy * ai[TXZ];
            ha[TYZ] = DetTyz * DBy - APDy * ai[TYZ];
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
                for (int c = VX; c <= VZ; ++c) ha[c] = a.timg[(6 + c) * a.tstr + tcell];
            }
            CGFD_AUX_UPDATE(1, ha, ai)
        }
        const bool inz = z < a.nabs[4] || z >= g.nk - a.nabs[5];
        if (inz) {
            const int iz = z < a.nabs[4] ? z : z - (g.nk - a.nabs[5]) + a.nabs[4];
            const long aidx = ((long)iz * g.nj + y) * row + x;
            R ai[NF], ha[NF];
#pragma unroll
            for (int c = 0; c < NF; ++c) ai[c] = a.ain[2][c * a.astr[2] + aidx];
#pragma unroll
            for (int c = 0; c < NF; ++c) hv[c] -= ai[c] / Bz;
            ha[VX] = DztVx * DBz - APDz * ai[VX];
            ha[VY] = DztVy * DBz - APDz * ai[VY];
            ha[VZ] = DztVz * DBz - APDz * ai[VZ];
            ha[TXX] = DztTxx * DBz - APDz * ai[TXX];
            ha[TYY] = DztTyy * DBz - APDz * ai[TYY];
            ha[TZZ] = DztTzz * DBz - APDz * ai[TZZ];
            ha[TXY] = DztTxy * DBz - APDz * ai[TXY];
            ha[TXZ] = DztTxz * DBz - APDz * ai[TXZ];
            ha[TYZ] = DztTyz * DBz - APDz * ai[TYZ];
            CGFD_AUX_UPDATE(2, ha, ai)
        }
#undef CGFD_AUX_UPDATE

        // ---- CalTIMG override of the velocity rates on the top three rows -----
        if (top >= 0) {
#pragma unroll
            for (int c = VX; c <= VZ; ++c) hv[c] = a.timg[c * a.tstr + tcell];
        }

        // ---- LoadForce then LoadMoment, in source order ------------------------
        if (e >= 0) {
            int4 en = a.sent[e];
            while (en.w == col && en.x == z) {
                const R *cf = a.scoef + (long)e * 8;
                if (en.y == 0) {
                    const R stf = a.fstf[(long)en.z * a.nstf_f + a.tindex];
                    const R A = cf[0];
                    hv[VX] += stf * cf[1] * A;
                    hv[VY] += stf * cf[2] * A;
                    hv[VZ] += stf * cf[3] * A;
                } else {
                    const R stf = a.mstf[(long)en.z * a.nstf_m + a.tindex];
                    R A = (R)1;
                    A = stf * A / cf[0];
                    hv[TXX] -= cf[1] * A;
                    hv[TYY] -= cf[2] * A;
                    hv[TZZ] -= cf[3] * A;
                    hv[TXY] -= cf[4] * A;
                    hv[TXZ] -= cf[5] * A;
                    hv[TYZ] -= cf[6] * A;
                }
                en = a.sent[++e];
            }
        }

        // ---- RK4 update (Iteration*, calculate.cu:6330-6852) -------------------
#pragma unroll
        for (int c = 0; c < NF; ++c) {
            const long q = c * fs + p;
            if (ST == 0) {
                const R f0 = w[c][CI]; // the stage input is FW itself
                a.wout[q] = f0 + a.alpha * hv[c];
                a.tw[q] = f0 + a.beta * hv[c];
            } else if (ST < 3) {
                a.wout[q] = a.fw[q] + a.alpha * hv[c];
                a.tw[q] = a.tw[q] + a.beta * hv[c];
            } else {
                const R wn = a.tw[q] + a.beta * hv[c];
                a.wout[q] = wn;
                if (c < 3 && top == 2 && a.pgv) {
                    R *pv = a.pgv + c * a.cstr + cxy; // PGV maps share the [nj][nxp] slot size of cov
                    *pv = rmax(rabs(wn), rabs(*pv));
                }
            }
        }

        // ---- slide the z window ---------------------------------------------------
#pragma unroll
        for (int c = 0; c < NF; ++c) {
            w[c][0] = w[c][1];
            w[c][1] = w[c][2];
            w[c][2] = w[c][3];
            w[c][3] = w[c][4];
        }
    }
}

// ---------------------------------------------------------------------------
// traction image, top three rows (CalTIMG, calculate.cu:5428-5666)
// ---------------------------------------------------------------------------
template <typename R, int SX, int SZ>
__global__ void __launch_bounds__(256) timg_kernel(TimgArgs<R> a)
{
    const Geom g = a.g;
    const int x = blockIdx.x * 64 + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= g.ni || y >= g.nj) return;
    const long fs = g.fs, row = g.nxp, pl = g.plane;
    const R *pmx = a.pml;
    const R *pmy = a.pml + 3 * g.ni;
    const R APDx = pmx[x], Bx = pmx[g.ni + x], DBx = pmx[2 * g.ni + x];
    const R APDy = pmy[y], By = pmy[g.nj + y], DBy = pmy[2 * g.nj + y];
    const bool inx = x < a.nabs[0] || x >= g.ni - a.nabs[1];
    const int ix = x < a.nabs[0] ? x : x - (g.ni - a.nabs[1]) + a.nabs[0];
    const bool iny = y < a.nabs[2] || y >= g.nj - a.nabs[3];
    const int iy = y < a.nabs[2] ? y : y - (g.nj - a.nabs[3]) + a.nabs[2];
    const double hx = SX > 0 ? a.h : -a.h;
    const double hz = SZ > 0 ? a.h : -a.h;
    // momentum rows: (Txx,Txy,Txz) -> Vx, (Txy,Tyy,Tyz) -> Vy, (Txz,Tyz,Tzz) -> Vz
    constexpr int trow[3][3] = {{TXX, TXY, TXZ}, {TXY, TYY, TYZ}, {TXZ, TYZ, TZZ}};
    const R *jac = a.met + JAC * fs;

#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const int z = g.nk - 3 + r;
        const long G = gidx(g, x, y, z);
        const R rhojac = (R)(1.0 / (double)a.med[2 * fs + G] / (double)jac[G]);
        const long tcell = ((long)r * g.nj + y) * row + x;
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            const R *T0 = a.win + trow[m][0] * fs;
            const R *T1 = a.win + trow[m][1] * fs;
            const R *T2 = a.win + trow[m][2] * fs;
            R vx[7], vy[7], vz[7];
#pragma unroll
            for (int i = 0; i < 7; ++i) {
                vx[i] = vy[i] = 0;
                if ((SX > 0 && i >= 2) || (SX < 0 && i <= 4)) {
                    long q = G + (i - 3);
                    vx[i] = jac[q] * (a.met[XIX * fs + q] * T0[q] + a.met[XIY * fs + q] * T1[q] +
                                      a.met[XIZ * fs + q] * T2[q]);
                    q = G + (i - 3) * row;
                    vy[i] = jac[q] * (a.met[ETX * fs + q] * T0[q] + a.met[ETY * fs + q] * T1[q] +
                                      a.met[ETZ * fs + q] * T2[q]);
                }
                const long q = G + (i - 3) * pl;
                vz[i] = jac[q] * (a.met[ZTX * fs + q] * T0[q] + a.met[ZTY * fs + q] * T1[q] +
                                  a.met[ZTZ * fs + q] * T2[q]);
            }
            // odd images above the surface, zero traction on it
            const int s = 3 + (2 - r);
#pragma unroll
            for (int j = 1; j <= 3; ++j)
                if (j <= 3 - (2 - r)) vz[s + j] = (R)0 - vz[s - j];
            vz[s] = 0;

            const R DxTx = (R)((double)rhojac * (drp_sum_strf<R, SX>(vx) / hx));
            const R DyTy = (R)((double)rhojac * (drp_sum_strf<R, SX>(vy) / hx));
            const R DzTz = (R)((double)rhojac * (drp_sum_strf<R, SZ>(vz) / hz));
            R hv = (Bx == (R)1 && By == (R)1) ? DxTx + DyTy + DzTz : DxTx / Bx + DyTy / By + DzTz;
            if (inx) {
                const R av = a.ain[0][m * a.astr[0] + ((long)z * g.nj + y) * a.npml[0] + ix];
                hv -= av / Bx;
                a.timg[(3 + m) * a.tstr + tcell] = DBx * DxTx - APDx * av;
            }
            if (iny) {
                const R av = a.ain[1][m * a.astr[1] + ((long)z * a.npml[1] + iy) * row + x];
                hv -= av / By;
                a.timg[(6 + m) * a.tstr + tcell] = DBy * DyTy - APDy * av;
            }
            a.timg[m * a.tstr + tcell] = hv;
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
    const R miu = med[2 * fs + G] * med[fs + G] * med[fs + G];
    const R lam2mu = med[2 * fs + G] * med[G] * med[G];
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
        A = (R)1 / (A * med[2 * g.fs + G]) * d;
        if (z == g.nk - 1) A = A * (R)2;
        cf[0] = A;
        for (int k = 0; k < 3; ++k) cf[1 + k] = fvec[3 * en.z + k];
    } else {
        // calculate.cu:6238-6240: the divisor h*h*h*jac
        cf[0] = hr * hr * hr * jac;
        for (int k = 0; k < 6; ++k) cf[1 + k] = mvec[6 * en.z + k];
    }
}

// receivers (PWpick, calculate.cu:2780-2828): all 9 components of FW
template <typename R>
__global__ void record_kernel(const R *fw, long fs, const long *rlin, int nrecv, R *seis, int nt, int t)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nrecv * NF) return;
    const int r = i / NF, c = i % NF;
    seis[((long)c * nrecv + r) * nt + t] = fw[c * fs + rlin[r]];
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <typename R, int SX, int SZ>
static void launch_stage_sxz(int st, const StageArgs<R> &a, dim3 grid, dim3 block, hipStream_t s)
{
    switch (st) {
    case 0: stage_kernel<R, SX, SZ, 0><<<grid, block, 0, s>>>(a); break;
    case 1: stage_kernel<R, SX, SZ, 1><<<grid, block, 0, s>>>(a); break;
    case 2: stage_kernel<R, SX, SZ, 2><<<grid, block, 0, s>>>(a); break;
    default: stage_kernel<R, SX, SZ, 3><<<grid, block, 0, s>>>(a); break;
    }
}

template <typename R>
void launch_stage(int sx, int sz, int st, const StageArgs<R> &a, int rows, int zchunks, hipStream_t s)
{
    dim3 block(64, rows);
    dim3 grid((a.g.ni + 63) / 64, (a.g.nj + rows - 1) / rows, zchunks);
    if (sx > 0 && sz > 0) launch_stage_sxz<R, 1, 1>(st, a, grid, block, s);
    else if (sx > 0) launch_stage_sxz<R, 1, -1>(st, a, grid, block, s);
    else if (sz > 0) launch_stage_sxz<R, -1, 1>(st, a, grid, block, s);
    else launch_stage_sxz<R, -1, -1>(st, a, grid, block, s);
}

template <typename R> void launch_timg(int sx, int sz, const TimgArgs<R> &a, int rows, hipStream_t s)
{
    dim3 block(64, rows);
    dim3 grid((a.g.ni + 63) / 64, (a.g.nj + rows - 1) / rows);
    if (sx > 0 && sz > 0) timg_kernel<R, 1, 1><<<grid, block, 0, s>>>(a);
    else if (sx > 0) timg_kernel<R, 1, -1><<<grid, block, 0, s>>>(a);
    else if (sz > 0) timg_kernel<R, -1, 1><<<grid, block, 0, s>>>(a);
    else timg_kernel<R, -1, -1><<<grid, block, 0, s>>>(a);
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
void launch_record(const R *fw, long fs, const long *rlin, int nrecv, R *seis, int nt, int t, hipStream_t s)
{
    if (nrecv <= 0) return;
    const int n = nrecv * NF;
    record_kernel<R><<<(n + 255) / 256, 256, 0, s>>>(fw, fs, rlin, nrecv, seis, nt, t);
}

#define CGFD_INSTANTIATE(R)                                                                                   \
    template void launch_stage<R>(int, int, int, const StageArgs<R> &, int, int, hipStream_t);              \
    template void launch_timg<R>(int, int, const TimgArgs<R> &, int, hipStream_t);                          \
    template void launch_velpd<R>(const Geom &, const R *, const R *, const R *, R *, long, hipStream_t);     \
    template void launch_to_padded<R>(const R *, R *, const Geom &, hipStream_t);                            \
    template void launch_from_padded<R>(const R *, R *, const Geom &, hipStream_t);                          \
    template void launch_src_prep<R>(const Geom &, const R *, const R *, const int4 *, const R *, const R *,  \
                                     R *, int, R, hipStream_t);                                               \
    template void launch_record<R>(const R *, long, const long *, int, R *, int, int, hipStream_t);

CGFD_INSTANTIATE(float)
CGFD_INSTANTIATE(double)

} // namespace cgfd
