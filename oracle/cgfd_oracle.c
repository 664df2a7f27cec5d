/*
 * cgfd_oracle.c -- CPU oracle for the CGFD3D time-stepping hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see cgfd_oracle.h).  A plain-C restatement of the
 * reference's kernels; every function cites the reference file:line it
 * follows.  Expressions are written in the same operand order and with the
 * same literal types as the reference source, and the file is compiled with
 * -ffp-contract=off, so each floating-point operation rounds as the
 * reference's host build would (FP32 storage, FP64 stencil accumulation
 * because the stencil coefficients are double literals, typenew.h:52-106).
 *
 * Structure: one call of orc_stage() performs the reference's RKite
 * (calculate.cu:1355-1599) for one RK stage: CalDiffCL + CalWaveCL fused per
 * cell (CalWaveCL reads the derivative arrays only at its own cell, so the
 * 27 derivatives are evaluated in registers instead of stored), then CalTIMG,
 * then LoadForce / LoadMoment, then Iteration{Begin,Inner,Final,FinalPV}.
 * The reference's stage-input copies mW/mA (wavesyn/abssyn,
 * calculate.cu:2956-3116) are aliases here: the stage input is read in full
 * before the update overwrites it.
 */
#include "cgfd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define LFD 3

/* one-sided DRP/opt MacCormack coefficients, typenew.h:52-56 */
#define DA_1 (-0.30874)
#define DA0 (-0.6326)
#define DA1 1.23300
#define DA2 (-0.3334)
#define DA3 0.04168
/* Gottlieb 2-4 MacCormack, typenew.h:76-79 */
#define M24A0 (-1.1666666666666667)
#define M24A1 1.3333333333333333
#define M24A2 (-0.16666666666666666)

/* Planted mistakes.  The pin tests (tests/test_oracle_pins.py) compile extra
 * copies of this file with -DORC_PLANT=n, each carrying one plausible
 * misreading of the reference, to show that some pin fails on it.  The
 * oracle proper is built with ORC_PLANT = 0, where every block below
 * compiles to the reference's expression. */
#ifndef ORC_PLANT
#define ORC_PLANT 0
#endif
#define PLANT(n) (ORC_PLANT == (n))

#define OABS(a) ((a) > 0 ? (a) : -(a))        /* typenew.h:44 */
#define OMAX(a, b) ((a) >= (b) ? (a) : (b))   /* typenew.h:45 */

enum { VX, VY, VZ, TXX, TYY, TZZ, TXY, TXZ, TYZ, NF };
enum { XIX, XIY, XIZ, ETX, ETY, ETZ, ZTX, ZTY, ZTZ, JAC, NDRV };

struct orc {
    int ni, nj, nk, NX, NY, NZ;
    long N;
    real dh, dt;
    int nt;
    int currT, stage;
    real *drv[NDRV];
    real *alpha, *beta, *rho;
    int nabs[6];
    real *APD[3], *Bt[3], *DB[3];
    real *covx, *covy; /* [NX][NY][9] */
    real *FW[NF], *W[NF], *tW[NF], *hW[NF];
    long asz[3];
    real *FA[3][NF], *A[3][NF], *tA[3][NF], *hA[3][NF];
    int nfrc, nmnt, nstf_f, nstf_m;
    int *floc, *mloc;
    real *fvec, *mvec, *fstf, *mstf;
    int nfoc, fnt;
    real fdt;
    int *rmloc;
    real *rmser; /* [6][nfoc*fnt]: mxx myy mzz mxy mxz myz series */
    int nrecv;
    int *rloc;
    real *seis;
    int pgv_on;
    real *pgv[3];
    int conindex; /* ConIndex: the curvilinear contraction applies at idz >= conindex */
    int simplecar; /* the SimpleCar build: xi/eta metrics taken as the identity */
};

static void *zalloc(size_t n) { return calloc(n ? n : 1, 1); }

/* RK4 coefficients as the reference stores them (calculate.cu:12-13) */
static const real RK4A[3] = {0.5, 0.5, 1.0};
#if PLANT(16) /* second and fourth RK weights exchanged */
static const real RK4B[4] = {1.0 / 6.0, 1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0};
#else
static const real RK4B[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
#endif

/* ---- stencils (typenew.h:58-60, 82-88) ---------------------------------- */
static inline double drp_fwd(const real *v, long i, long inc, double dh)
{
    return (DA_1 * v[i - 1 * inc] + DA0 * v[i] + DA1 * v[i + 1 * inc] + DA2 * v[i + 2 * inc] +
            DA3 * v[i + 3 * inc]) / dh;
}
#if PLANT(15) /* DA2 and DA3 exchanged in the backward sum */
static inline double drp_bwd(const real *v, long i, long inc, double dh)
{
    return (DA2 * v[i + 3 * inc] + DA3 * v[i + 2 * inc] + DA1 * v[i + 1 * inc] + DA0 * v[i] +
            DA_1 * v[i - 1 * inc]) / dh;
}
#else
static inline double drp_bwd(const real *v, long i, long inc, double dh)
{
    return (DA3 * v[i + 3 * inc] + DA2 * v[i + 2 * inc] + DA1 * v[i + 1 * inc] + DA0 * v[i] +
            DA_1 * v[i - 1 * inc]) / dh;
}
#endif
static inline real drpfd(const real *v, long i, real dh, long inc)
{
    return dh > 0 ? (real)drp_fwd(v, i, inc, dh) : (real)drp_bwd(v, i, inc, dh);
}
static inline real m24fd(const real *v, long i, real dh, long inc)
{
#if PLANT(14) /* backward 2-4 closure divided by |h| */
    return dh > 0 ? (real)((M24A0 * v[i] + M24A1 * v[i + 1 * inc] + M24A2 * v[i + 2 * inc]) / dh)
                  : (real)((M24A2 * v[i + 2 * inc] + M24A1 * v[i + 1 * inc] + M24A0 * v[i]) / -dh);
#else
    return dh > 0 ? (real)((M24A0 * v[i] + M24A1 * v[i + 1 * inc] + M24A2 * v[i + 2 * inc]) / dh)
                  : (real)((M24A2 * v[i + 2 * inc] + M24A1 * v[i + 1 * inc] + M24A0 * v[i]) / dh);
#endif
}
static inline real m22fd(const real *v, long i, real dh, long inc)
{
    if (dh > 0)
#if PLANT(13) /* forward two-point difference with its sign flipped */
        return (v[i] - v[i + 1 * inc]) / dh;
#else
        return (v[i + 1 * inc] - v[i]) / dh;
#endif
    return (real)(-1.0 * (v[i] - v[i + 1 * inc]) / dh);
}

/* ---- lifecycle ------------------------------------------------------------ */
orc_t *orc_create(int ni, int nj, int nk, real dh, real dt, int nt)
{
    if (ni < 1 || nj < 1 || nk < 4 || nt < 1)
        return NULL;
    orc_t *o = (orc_t *)zalloc(sizeof(orc_t));
    o->ni = ni; o->nj = nj; o->nk = nk;
    o->NX = ni + 2 * LFD; o->NY = nj + 2 * LFD; o->NZ = nk + 2 * LFD;
    o->N = (long)o->NX * o->NY * o->NZ;
    o->dh = dh; o->dt = dt; o->nt = nt;
    o->conindex = LFD; /* hybrid_grid = 0: ConIndex = nk1 (gridmesh.cpp:441, simulation.cc:113) */
    for (int k = 0; k < NDRV; k++) o->drv[k] = (real *)zalloc(o->N * sizeof(real));
    o->alpha = (real *)zalloc(o->N * sizeof(real));
    o->beta = (real *)zalloc(o->N * sizeof(real));
    o->rho = (real *)zalloc(o->N * sizeof(real));
    int len[3] = {o->NX, o->NY, o->NZ};
    for (int a = 0; a < 3; a++) {
        o->APD[a] = (real *)zalloc(len[a] * sizeof(real));
        o->Bt[a] = (real *)zalloc(len[a] * sizeof(real));
        o->DB[a] = (real *)zalloc(len[a] * sizeof(real));
        for (int i = 0; i < len[a]; i++) o->Bt[a][i] = 1;
    }
    o->covx = (real *)zalloc((size_t)o->NX * o->NY * 9 * sizeof(real));
    o->covy = (real *)zalloc((size_t)o->NX * o->NY * 9 * sizeof(real));
    for (int c = 0; c < NF; c++) {
        o->FW[c] = (real *)zalloc(o->N * sizeof(real));
        o->W[c] = (real *)zalloc(o->N * sizeof(real));
        o->tW[c] = (real *)zalloc(o->N * sizeof(real));
        o->hW[c] = (real *)zalloc(o->N * sizeof(real));
    }
    for (int a = 0; a < 3; a++) o->pgv[a] = (real *)zalloc((size_t)o->NX * o->NY * sizeof(real));
    return o;
}

static void free_aux(orc_t *o)
{
    for (int a = 0; a < 3; a++)
        for (int c = 0; c < NF; c++) {
            free(o->FA[a][c]); free(o->A[a][c]); free(o->tA[a][c]); free(o->hA[a][c]);
            o->FA[a][c] = o->A[a][c] = o->tA[a][c] = o->hA[a][c] = NULL;
        }
}

void orc_destroy(orc_t *o)
{
    if (!o) return;
    for (int k = 0; k < NDRV; k++) free(o->drv[k]);
    free(o->alpha); free(o->beta); free(o->rho);
    for (int a = 0; a < 3; a++) { free(o->APD[a]); free(o->Bt[a]); free(o->DB[a]); free(o->pgv[a]); }
    free(o->covx); free(o->covy);
    for (int c = 0; c < NF; c++) { free(o->FW[c]); free(o->W[c]); free(o->tW[c]); free(o->hW[c]); }
    free_aux(o);
    free(o->floc); free(o->mloc); free(o->fvec); free(o->mvec); free(o->fstf); free(o->mstf);
    free(o->rmloc); free(o->rmser);
    free(o->rloc); free(o->seis);
    free(o);
}

int orc_set_metrics(orc_t *o, const real *const *drv)
{
    for (int k = 0; k < NDRV; k++) memcpy(o->drv[k], drv[k], o->N * sizeof(real));
    return 0;
}

int orc_set_media(orc_t *o, const real *alpha, const real *beta, const real *rho)
{
    memcpy(o->alpha, alpha, o->N * sizeof(real));
    memcpy(o->beta, beta, o->N * sizeof(real));
    memcpy(o->rho, rho, o->N * sizeof(real));
    return 0;
}

int orc_set_pml(orc_t *o, const int *nabs, const real *const *tabs)
{
    int len[3] = {o->NX, o->NY, o->NZ};
    int nv[3] = {o->ni, o->nj, o->nk};
    for (int a = 0; a < 3; a++)
        if (nabs[2 * a] < 0 || nabs[2 * a + 1] < 0 || nabs[2 * a] + nabs[2 * a + 1] >= nv[a])
            return -1; /* absorb.cpp:151-162 */
    memcpy(o->nabs, nabs, 6 * sizeof(int));
    for (int a = 0; a < 3; a++) {
        memcpy(o->APD[a], tabs[a], len[a] * sizeof(real));
        memcpy(o->Bt[a], tabs[3 + a], len[a] * sizeof(real));
        memcpy(o->DB[a], tabs[6 + a], len[a] * sizeof(real));
    }
    /* packed auxiliary sizes, calculate.cu:325-327 */
    free_aux(o);
    o->asz[0] = (long)(nabs[0] + nabs[1]) * o->NY * o->NZ;
    o->asz[1] = (long)(nabs[2] + nabs[3]) * o->NX * o->NZ;
    o->asz[2] = (long)(nabs[4] + nabs[5]) * o->NX * o->NY;
    for (int a = 0; a < 3; a++)
        for (int c = 0; c < NF; c++) {
            o->FA[a][c] = (real *)zalloc(o->asz[a] * sizeof(real));
            o->A[a][c] = (real *)zalloc(o->asz[a] * sizeof(real));
            o->tA[a][c] = (real *)zalloc(o->asz[a] * sizeof(real));
            o->hA[a][c] = (real *)zalloc(o->asz[a] * sizeof(real));
        }
    return 0;
}

static int copy_src(orc_t *o, int n, const int *locs, const real *vec, int nv, const real *stf, int nstf,
                    int **ploc, real **pvec, real **pstf)
{
    for (int i = 0; i < n; i++) {
        int x = locs[3 * i], y = locs[3 * i + 1], z = locs[3 * i + 2];
        if (x < LFD || x >= o->ni + LFD || y < LFD || y >= o->nj + LFD || z < LFD || z >= o->nk + LFD)
            return -1;
    }
    free(*ploc); free(*pvec); free(*pstf);
    *ploc = (int *)zalloc((size_t)3 * n * sizeof(int));
    *pvec = (real *)zalloc((size_t)nv * n * sizeof(real));
    *pstf = (real *)zalloc((size_t)n * nstf * sizeof(real));
    memcpy(*ploc, locs, (size_t)3 * n * sizeof(int));
    memcpy(*pvec, vec, (size_t)nv * n * sizeof(real));
    memcpy(*pstf, stf, (size_t)n * nstf * sizeof(real));
    return 0;
}

int orc_set_forces(orc_t *o, int n, const int *locs, const real *vec, const real *stf, int nstf)
{
    if (n > 0 && nstf < 2 * o->nt + 1) return -1;
    if (copy_src(o, n, locs, vec, 3, stf, nstf, &o->floc, &o->fvec, &o->fstf)) return -1;
    o->nfrc = n; o->nstf_f = nstf;
    return 0;
}

int orc_set_moments(orc_t *o, int n, const int *locs, const real *vec, const real *stf, int nstf)
{
    if (n > 0 && nstf < 2 * o->nt + 1) return -1;
    if (copy_src(o, n, locs, vec, 6, stf, nstf, &o->mloc, &o->mvec, &o->mstf)) return -1;
    o->nmnt = n; o->nstf_m = nstf;
    return 0;
}

int orc_set_focal(orc_t *o, int n, const int *locs, const real *series, int fnt, real fdt)
{
    if (n < 0 || (n > 0 && (fnt < 1 || !(fdt > 0)))) return -1;
    for (int i = 0; i < n; i++) {
        int x = locs[3 * i], y = locs[3 * i + 1], z = locs[3 * i + 2];
        if (x < LFD || x >= o->ni + LFD || y < LFD || y >= o->nj + LFD || z < LFD || z >= o->nk + LFD)
            return -1;
    }
    free(o->rmloc); free(o->rmser);
    o->rmloc = (int *)zalloc((size_t)3 * n * sizeof(int));
    o->rmser = (real *)zalloc((size_t)6 * n * (n > 0 ? fnt : 1) * sizeof(real));
    memcpy(o->rmloc, locs, (size_t)3 * n * sizeof(int));
    if (n > 0) memcpy(o->rmser, series, (size_t)6 * n * fnt * sizeof(real));
    o->nfoc = n; o->fnt = fnt; o->fdt = fdt;
    return 0;
}

int orc_set_receivers(orc_t *o, int n, const int *locs)
{
    for (int i = 0; i < n; i++) {
        int x = locs[3 * i], y = locs[3 * i + 1], z = locs[3 * i + 2];
        if (x < LFD || x >= o->ni + LFD || y < LFD || y >= o->nj + LFD || z < LFD || z >= o->nk + LFD)
            return -1;
    }
    free(o->rloc); free(o->seis);
    o->nrecv = n;
    o->rloc = (int *)zalloc((size_t)3 * n * sizeof(int));
    memcpy(o->rloc, locs, (size_t)3 * n * sizeof(int));
    o->seis = (real *)zalloc((size_t)9 * n * o->nt * sizeof(real));
    return 0;
}

int orc_set_pgv(orc_t *o, int on) { o->pgv_on = on ? 1 : 0; return 0; }

/* hybrid_grid = 1: ConIndex from the hybrid mesh (gridmesh.cpp:346); the
 * three surface rows stay curvilinear */
int orc_set_conindex(orc_t *o, int conindex)
{
    if (conindex < LFD || conindex > o->nk + LFD - 3) return -1;
    o->conindex = conindex;
    return 0;
}

/* The reference's SimpleCar build (its Makefile default): the xi/eta metric
 * arrays are never allocated and CalDiffmix/CalWavemix/CalTIMGmix replace
 * CalDiff/CalWave/CalTIMG (calculate.cu:1471-1480, 1509-1511), while
 * VelPDcoeff takes e11 = e22 = 1 and the other xi/eta entries 0
 * (calculate.cu:3362-3375). */
int orc_set_simplecar(orc_t *o, int on)
{
    o->simplecar = on ? 1 : 0;
    return 0;
}

/* ---- VelPDcoeff, calculate.cu:3320-3466 --------------------------------- */
static void matinv3(real *A) /* calculate.cu:3444-3466, Gauss-Jordan without pivoting */
{
    int n = 3;
    real con;
    for (int i = 0; i < n; i++) {
        con = A[i * n + i];
        A[i * n + i] = 1;
        for (int j = 0; j < n; j++) A[i * n + j] = A[i * n + j] / con;
        for (int j = 0; j < n; j++)
            if (j != i) {
                con = A[j * n + i];
                A[j * n + i] = 0;
                for (int k = 0; k < n; k++) A[j * n + k] = A[j * n + k] - A[i * n + k] * con;
            }
    }
}

static void matmul3(const real *A, const real *B, real *C) /* calculate.cu:3430-3442 */
{
    int n = 3;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            C[i * n + j] = 0.0;
            for (int k = 0; k < n; k++) C[i * n + j] = C[i * n + j] + A[i * n + k] * B[k * n + j];
        }
}

static void velpdcoeff(orc_t *o)
{
    const int NY = o->NY, NZ = o->NZ;
#pragma omp parallel for schedule(static)
    for (int idx = LFD; idx < o->ni + LFD; idx++)
        for (int idy = LFD; idy < o->nj + LFD; idy++) {
            int idz = o->nk + LFD - 1;
            long G = ((long)idx * NY + idy) * NZ + idz;
#if PLANT(4) /* PML ratio inverted */
            real Bzx = o->Bt[0][idx] / o->Bt[2][idz];
#else
            real Bzx = o->Bt[2][idz] / o->Bt[0][idx];
#endif
            real Bzy = o->Bt[2][idz] / o->Bt[1][idy];
            real miu = o->rho[G] * o->beta[G] * o->beta[G];
            real lam2mu = o->rho[G] * o->alpha[G] * o->alpha[G];
            real lambda = lam2mu - 2.0 * miu;
            real e11 = o->drv[XIX][G], e12 = o->drv[XIY][G], e13 = o->drv[XIZ][G];
            real e21 = o->drv[ETX][G], e22 = o->drv[ETY][G], e23 = o->drv[ETZ][G];
            if (o->simplecar) { /* calculate.cu:3369-3375 */
                e11 = 1; e12 = 0; e13 = 0;
                e21 = 0; e22 = 1; e23 = 0;
            }
            real e31 = o->drv[ZTX][G], e32 = o->drv[ZTY][G], e33 = o->drv[ZTZ][G];
            real A[9], B[9], C[9], tmp[9];
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
#if PLANT(3) /* entry (0,1) of the xi block transposed */
            B[1] = lambda * e32 * e11 + miu * e31 * e12;
#else
            B[1] = lambda * e31 * e12 + miu * e32 * e11;
#endif
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
            long xi = ((long)idx * NY + idy) * 9;
            matmul3(A, B, tmp);
            for (int i = 0; i < 9; i++) o->covx[xi + i] = -1 * tmp[i] * Bzx;
            matmul3(A, C, tmp);
            for (int i = 0; i < 9; i++) o->covy[xi + i] = -1 * tmp[i] * Bzy;
        }
}

int orc_init(orc_t *o)
{
    for (int c = 0; c < NF; c++) {
        memset(o->FW[c], 0, o->N * sizeof(real));
        memset(o->W[c], 0, o->N * sizeof(real));
        memset(o->tW[c], 0, o->N * sizeof(real));
        memset(o->hW[c], 0, o->N * sizeof(real));
    }
    for (int a = 0; a < 3; a++) {
        for (int c = 0; c < NF; c++) {
            if (!o->FA[a][c]) continue;
            memset(o->FA[a][c], 0, o->asz[a] * sizeof(real));
            memset(o->A[a][c], 0, o->asz[a] * sizeof(real));
            memset(o->tA[a][c], 0, o->asz[a] * sizeof(real));
            memset(o->hA[a][c], 0, o->asz[a] * sizeof(real));
        }
        memset(o->pgv[a], 0, (size_t)o->NX * o->NY * sizeof(real));
    }
    if (o->seis) memset(o->seis, 0, (size_t)9 * o->nrecv * o->nt * sizeof(real));
    o->currT = 0;
    o->stage = 0;
    velpdcoeff(o);
    return 0;
}

/* ---- RHS: CalDiffCL (calculate.cu:4623-4801) + CalWaveCL (4804-5118) ------ */
static void rhs(orc_t *o, real *const *mW, real *const mA[3][NF], int Xvec, int Yvec, int Zvec)
{
    const int NY = o->NY, NZ = o->NZ, nk = o->nk;
    const real steph = o->dh;
    const real xstep = steph * Xvec, ystep = steph * Yvec, zstep = steph * Zvec;
    const long xinc = (long)Xvec * NY * NZ, yinc = (long)Yvec * NZ, zinc = Zvec * 1;
    const int *nabs = o->nabs;

#pragma omp parallel for schedule(static)
    for (int idx = LFD; idx < o->ni + LFD; idx++)
        for (int idy = LFD; idy < o->nj + LFD; idy++)
            for (int idz = LFD; idz < nk + LFD; idz++) {
                long G = ((long)idx * NY + idy) * NZ + idz;
                long xiaoI = ((long)idx * NY + idy) * 9;
                const real *Vx = mW[VX], *Vy = mW[VY], *Vz = mW[VZ];
                const real *Txx = mW[TXX], *Tyy = mW[TYY], *Tzz = mW[TZZ];
                const real *Txy = mW[TXY], *Txz = mW[TXZ], *Tyz = mW[TYZ];
                /* CalDiffCL: 27 one-sided derivatives */
                real DxTxx = drpfd(Txx, G, xstep, xinc), DxTyy = drpfd(Tyy, G, xstep, xinc);
                real DxTzz = drpfd(Tzz, G, xstep, xinc), DxTxy = drpfd(Txy, G, xstep, xinc);
                real DxTxz = drpfd(Txz, G, xstep, xinc), DxTyz = drpfd(Tyz, G, xstep, xinc);
                real DxVx = drpfd(Vx, G, xstep, xinc), DxVy = drpfd(Vy, G, xstep, xinc);
                real DxVz = drpfd(Vz, G, xstep, xinc);
                real DyTxx = drpfd(Txx, G, ystep, yinc), DyTyy = drpfd(Tyy, G, ystep, yinc);
                real DyTzz = drpfd(Tzz, G, ystep, yinc), DyTxy = drpfd(Txy, G, ystep, yinc);
                real DyTxz = drpfd(Txz, G, ystep, yinc), DyTyz = drpfd(Tyz, G, ystep, yinc);
                real DyVx = drpfd(Vx, G, ystep, yinc), DyVy = drpfd(Vy, G, ystep, yinc);
                real DyVz = drpfd(Vz, G, ystep, yinc);
                real DzTxx = drpfd(Txx, G, zstep, zinc), DzTyy = drpfd(Tyy, G, zstep, zinc);
                real DzTzz = drpfd(Tzz, G, zstep, zinc), DzTxy = drpfd(Txy, G, zstep, zinc);
                real DzTxz = drpfd(Txz, G, zstep, zinc), DzTyz = drpfd(Tyz, G, zstep, zinc);
                real DzVx, DzVy, DzVz;
                const real *CoVx = o->covx, *CoVy = o->covy;
                /* CondFreeVLOW velocity z-derivative closure, calculate.cu:4736-4769 */
                if (idz == nk + LFD - 1) {
                    DzVx = CoVx[xiaoI + 0] * DxVx + CoVx[xiaoI + 1] * DxVy + CoVx[xiaoI + 2] * DxVz +
                           CoVy[xiaoI + 0] * DyVx + CoVy[xiaoI + 1] * DyVy + CoVy[xiaoI + 2] * DyVz;
                    DzVy = CoVx[xiaoI + 3] * DxVx + CoVx[xiaoI + 4] * DxVy + CoVx[xiaoI + 5] * DxVz +
                           CoVy[xiaoI + 3] * DyVx + CoVy[xiaoI + 4] * DyVy + CoVy[xiaoI + 5] * DyVz;
                    DzVz = CoVx[xiaoI + 6] * DxVx + CoVx[xiaoI + 7] * DxVy + CoVx[xiaoI + 8] * DxVz +
                           CoVy[xiaoI + 6] * DyVx + CoVy[xiaoI + 7] * DyVy + CoVy[xiaoI + 8] * DyVz;
                } else if (idz == nk + LFD - 2) {
                    DzVx = m22fd(Vx, G, zstep, zinc);
                    DzVy = m22fd(Vy, G, zstep, zinc);
                    DzVz = m22fd(Vz, G, zstep, zinc);
                } else if (idz == nk + LFD - 3) {
                    DzVx = m24fd(Vx, G, zstep, zinc);
                    DzVy = m24fd(Vy, G, zstep, zinc);
                    DzVz = m24fd(Vz, G, zstep, zinc);
                } else {
                    DzVx = drpfd(Vx, G, zstep, zinc);
                    DzVy = drpfd(Vy, G, zstep, zinc);
                    DzVz = drpfd(Vz, G, zstep, zinc);
                }

                /* CalWaveCL: contravariant contraction, Hooke and momentum */
                real rho = o->rho[G];
                real miu = rho * o->beta[G] * o->beta[G];
                real lam2mu = rho * o->alpha[G] * o->alpha[G];
                real lambda = lam2mu - 2.0 * miu;
                rho = 1.0 / rho;
                /* CalWave (hybrid grid) below the conversion depth loads only xi_x,
                 * eta_y and zeta_z and uses a Cartesian form of the contraction
                 * (calculate.cu:3828-3956); at and above ConIndex, and everywhere
                 * without the hybrid grid (ConIndex = nk1), the curvilinear
                 * contraction of CalWaveCL (calculate.cu:4855-4930) */
                real xix = o->drv[XIX][G], ety = o->drv[ETY][G], ztz = o->drv[ZTZ][G];
                real xiy = 0, xiz = 0, etx = 0, etz = 0, ztx = 0, zty = 0;
                real DxiVx, DetVx, DztVx, DxiVy, DetVy, DztVy, DxiVz, DetVz, DztVz;
                real DxiTxx, DetTxx, DztTxx, DxiTyy, DetTyy, DztTyy, DxiTzz, DetTzz, DztTzz;
                real DxiTxy, DetTxy, DztTxy, DxiTxz, DetTxz, DztTxz, DxiTyz, DetTyz, DztTyz;
                if (o->simplecar && idz >= o->conindex) {
                    /* CalWavemix above ConIndex (calculate.cu:4339-4407): zeta
                     * metrics only */
                    ztx = o->drv[ZTX][G];
                    zty = o->drv[ZTY][G];
                    DxiVx = DxTxx * rho;
                    DetVx = DyTxy * rho;
                    DztVx = (ztx * DzTxx + zty * DzTxy + ztz * DzTxz) * rho;
                    DxiVy = DxTxy * rho;
                    DetVy = DyTyy * rho;
                    DztVy = (ztx * DzTxy + zty * DzTyy + ztz * DzTyz) * rho;
                    DxiVz = DxTxz * rho;
                    DetVz = DyTyz * rho;
                    DztVz = (ztx * DzTxz + zty * DzTyz + ztz * DzTzz) * rho;
                    DxiTxx = lam2mu * DxVx;
#if PLANT(18) /* lam2mu in place of lambda in the eta term of Txx (calculate.cu:4381) */
                    DetTxx = lam2mu * DyVy;
#else
                    DetTxx = lambda * DyVy;
#endif
                    DztTxx = lam2mu * ztx * DzVx + lambda * zty * DzVy + lambda * ztz * DzVz;
                    DxiTyy = lambda * DxVx;
                    DetTyy = lam2mu * DyVy;
                    DztTyy = lambda * ztx * DzVx + lam2mu * zty * DzVy + lambda * ztz * DzVz;
                    DxiTzz = lambda * DxVx;
                    DetTzz = lambda * DyVy;
                    DztTzz = lambda * ztx * DzVx + lambda * zty * DzVy + lam2mu * ztz * DzVz;
                    DxiTxy = DxVy * miu;
                    DetTxy = DyVx * miu;
                    DztTxy = (zty * DzVx + ztx * DzVy) * miu;
                    DxiTxz = DxVz * miu;
                    DetTxz = 0.0;
                    DztTxz = (ztz * DzVx + ztx * DzVz) * miu;
                    DxiTyz = 0.0;
                    DetTyz = DyVz * miu;
                    DztTyz = (ztz * DzVy + zty * DzVz) * miu;
                } else if (o->simplecar) {
                    /* CalWavemix below ConIndex (calculate.cu:4409-4456) */
                    DxiVx = rho * DxTxx;
                    DetVx = rho * DyTxy;
                    DztVx = ztz * DzTxz * rho;
                    DxiVy = rho * DxTxy;
                    DetVy = rho * DyTyy;
                    DztVy = ztz * DzTyz * rho;
                    DxiVz = rho * DxTxz;
                    DetVz = rho * DyTyz;
                    DztVz = ztz * DzTzz * rho;
                    DxiTxx = lam2mu * DxVx;
                    DetTxx = lambda * DyVy;
                    DztTxx = lambda * ztz * DzVz;
                    DxiTyy = lambda * DxVx;
                    DetTyy = lam2mu * DyVy;
                    DztTyy = lambda * ztz * DzVz;
                    DxiTzz = lambda * DxVx;
                    DetTzz = lambda * DyVy;
                    DztTzz = lam2mu * ztz * DzVz;
                    DxiTxy = miu * DxVy;
                    DetTxy = miu * DyVx;
                    DztTxy = 0.0;
                    DxiTxz = miu * DxVz;
                    DetTxz = 0.0;
                    DztTxz = ztz * DzVx * miu;
                    DxiTyz = 0.0;
                    DetTyz = miu * DyVz;
                    DztTyz = ztz * DzVy * miu;
                } else if (idz >= o->conindex) {
                    xiy = o->drv[XIY][G];
                    xiz = o->drv[XIZ][G];
                    etx = o->drv[ETX][G];
                    etz = o->drv[ETZ][G];
                    ztx = o->drv[ZTX][G];
                    zty = o->drv[ZTY][G];

                    DxiVx = (xix * DxTxx + xiy * DxTxy + xiz * DxTxz) * rho;
                    DetVx = (etx * DyTxx + ety * DyTxy + etz * DyTxz) * rho;
    #if PLANT(1) /* zeta_x and zeta_y swapped in the zeta term of the Vx rate */
                    DztVx = (zty * DzTxx + ztx * DzTxy + ztz * DzTxz) * rho;
    #else
                    DztVx = (ztx * DzTxx + zty * DzTxy + ztz * DzTxz) * rho;
    #endif
                    DxiVy = (xix * DxTxy + xiy * DxTyy + xiz * DxTyz) * rho;
                    DetVy = (etx * DyTxy + ety * DyTyy + etz * DyTyz) * rho;
                    DztVy = (ztx * DzTxy + zty * DzTyy + ztz * DzTyz) * rho;
                    DxiVz = (xix * DxTxz + xiy * DxTyz + xiz * DxTzz) * rho;
                    DetVz = (etx * DyTxz + ety * DyTyz + etz * DyTzz) * rho;
                    DztVz = (ztx * DzTxz + zty * DzTyz + ztz * DzTzz) * rho;

                    DxiTxx = lam2mu * xix * DxVx + lambda * xiy * DxVy + lambda * xiz * DxVz;
                    DetTxx = lam2mu * etx * DyVx + lambda * ety * DyVy + lambda * etz * DyVz;
                    DztTxx = lam2mu * ztx * DzVx + lambda * zty * DzVy + lambda * ztz * DzVz;
                    DxiTyy = lambda * xix * DxVx + lam2mu * xiy * DxVy + lambda * xiz * DxVz;
                    DetTyy = lambda * etx * DyVx + lam2mu * ety * DyVy + lambda * etz * DyVz;
                    DztTyy = lambda * ztx * DzVx + lam2mu * zty * DzVy + lambda * ztz * DzVz;
                    DxiTzz = lambda * xix * DxVx + lambda * xiy * DxVy + lam2mu * xiz * DxVz;
                    DetTzz = lambda * etx * DyVx + lambda * ety * DyVy + lam2mu * etz * DyVz;
                    DztTzz = lambda * ztx * DzVx + lambda * zty * DzVy + lam2mu * ztz * DzVz;
                    DxiTxy = (xiy * DxVx + xix * DxVy) * miu;
    #if PLANT(2) /* eta_x and eta_y swapped in the eta term of the Txy rate */
                    DetTxy = (etx * DyVx + ety * DyVy) * miu;
    #else
                    DetTxy = (ety * DyVx + etx * DyVy) * miu;
    #endif
                    DztTxy = (zty * DzVx + ztx * DzVy) * miu;
                    DxiTxz = (xiz * DxVx + xix * DxVz) * miu;
                    DetTxz = (etz * DyVx + etx * DyVz) * miu;
                    DztTxz = (ztz * DzVx + ztx * DzVz) * miu;
                    DxiTyz = (xiz * DxVy + xiy * DxVz) * miu;
                    DetTyz = (etz * DyVy + ety * DyVz) * miu;
                    DztTyz = (ztz * DzVy + zty * DzVz) * miu;
                } else {
                    DxiVx = rho * xix * DxTxx;
                    DetVx = rho * ety * DyTxy;
                    DztVx = ztz * DzTxz * rho;
                    DxiVy = rho * xix * DxTxy;
                    DetVy = rho * ety * DyTyy;
                    DztVy = ztz * DzTyz * rho;
                    DxiVz = rho * xix * DxTxz;
                    DetVz = rho * ety * DyTyz;
                    DztVz = ztz * DzTzz * rho;
                    DxiTxx = lam2mu * xix * DxVx;
                    DetTxx = lambda * ety * DyVy;
                    DztTxx = lambda * ztz * DzVz;
                    DxiTyy = lambda * xix * DxVx;
                    DetTyy = lam2mu * ety * DyVy;
                    DztTyy = lambda * ztz * DzVz;
                    DxiTzz = lambda * xix * DxVx;
                    DetTzz = lambda * ety * DyVy;
#if PLANT(17) /* lambda in place of lam2mu on the zz diagonal of the Cartesian branch (calculate.cu:3940) */
                    DztTzz = lambda * ztz * DzVz;
#else
                    DztTzz = lam2mu * ztz * DzVz;
#endif
                    DxiTxy = miu * xix * DxVy;
                    DetTxy = miu * ety * DyVx;
                    DztTxy = 0.0;
                    DxiTxz = miu * xix * DxVz;
                    DetTxz = 0.0;
                    DztTxz = ztz * DzVx * miu;
                    DxiTyz = 0.0;
                    DetTyz = miu * ety * DyVz;
                    DztTyz = ztz * DzVy * miu;
                }

                real APDx = o->APD[0][idx], APDy = o->APD[1][idy], APDz = o->APD[2][idz];
                real Bx = o->Bt[0][idx], By = o->Bt[1][idy], Bz = o->Bt[2][idz];
                real DBx = o->DB[0][idx], DBy = o->DB[1][idy], DBz = o->DB[2][idz];

                real h[NF];
                h[TXX] = DxiTxx / Bx + DetTxx / By + DztTxx / Bz;
                h[TYY] = DxiTyy / Bx + DetTyy / By + DztTyy / Bz;
                h[TZZ] = DxiTzz / Bx + DetTzz / By + DztTzz / Bz;
                h[TXY] = DxiTxy / Bx + DetTxy / By + DztTxy / Bz;
                h[TXZ] = DxiTxz / Bx + DetTxz / By + DztTxz / Bz;
                h[TYZ] = DxiTyz / Bx + DetTyz / By + DztTyz / Bz;
                h[VX] = DxiVx / Bx + DetVx / By + DztVx / Bz;
                h[VY] = DxiVy / Bx + DetVy / By + DztVy / Bz;
                h[VZ] = DxiVz / Bx + DetVz / By + DztVz / Bz;

                int surf = (idz == nk + LFD - 1);
                real DzVx1 = 0, DzVx2 = 0, DzVy1 = 0, DzVy2 = 0, DzVz1 = 0, DzVz2 = 0;
                if (surf) {
                    DzVx1 = CoVx[xiaoI + 0] * DxVx + CoVx[xiaoI + 1] * DxVy + CoVx[xiaoI + 2] * DxVz;
                    DzVx2 = CoVy[xiaoI + 0] * DyVx + CoVy[xiaoI + 1] * DyVy + CoVy[xiaoI + 2] * DyVz;
                    DzVy1 = CoVx[xiaoI + 3] * DxVx + CoVx[xiaoI + 4] * DxVy + CoVx[xiaoI + 5] * DxVz;
                    DzVy2 = CoVy[xiaoI + 3] * DyVx + CoVy[xiaoI + 4] * DyVy + CoVy[xiaoI + 5] * DyVz;
                    DzVz1 = CoVx[xiaoI + 6] * DxVx + CoVx[xiaoI + 7] * DxVy + CoVx[xiaoI + 8] * DxVz;
                    DzVz2 = CoVy[xiaoI + 6] * DyVx + CoVy[xiaoI + 7] * DyVy + CoVy[xiaoI + 8] * DyVz;
                }

                /* ADE CFS-PML, calculate.cu:4993-5107 */
                int tempIdx = idx;
                if (tempIdx <= nabs[0] + LFD - 1 || tempIdx >= o->ni + LFD - nabs[1]) {
                    long P = tempIdx < nabs[0] + LFD ? tempIdx - LFD : tempIdx - (o->ni + LFD - nabs[1]) + nabs[0];
                    P = (P * NY + idy) * NZ + idz;
                    real *const *Ax = mA[0];
                    real **hAx = o->hA[0];
                    for (int c = 0; c < NF; c++) h[c] -= Ax[c][P] / Bx;
                    hAx[VX][P] = DxiVx * DBx - APDx * Ax[VX][P];
                    hAx[VY][P] = DxiVy * DBx - APDx * Ax[VY][P];
                    hAx[VZ][P] = DxiVz * DBx - APDx * Ax[VZ][P];
                    hAx[TXX][P] = DxiTxx * DBx - APDx * Ax[TXX][P];
                    hAx[TYY][P] = DxiTyy * DBx - APDx * Ax[TYY][P];
                    hAx[TZZ][P] = DxiTzz * DBx - APDx * Ax[TZZ][P];
                    hAx[TXY][P] = DxiTxy * DBx - APDx * Ax[TXY][P];
                    hAx[TXZ][P] = DxiTxz * DBx - APDx * Ax[TXZ][P];
                    hAx[TYZ][P] = DxiTyz * DBx - APDx * Ax[TYZ][P];
                    if (surf) {
#if PLANT(10) /* surface term of the x auxiliary without the Bx factor */
                        hAx[TXX][P] += DBx * (lam2mu * ztx * DzVx1 + lambda * zty * DzVy1 + lambda * ztz * DzVz1);
#elif PLANT(11) /* the eta part of the closure routed into the x auxiliary */
                        hAx[TXX][P] += DBx * Bx * (lam2mu * ztx * DzVx2 + lambda * zty * DzVy2 + lambda * ztz * DzVz2);
#else
                        hAx[TXX][P] += DBx * Bx * (lam2mu * ztx * DzVx1 + lambda * zty * DzVy1 + lambda * ztz * DzVz1);
#endif
                        hAx[TYY][P] += DBx * Bx * (lambda * ztx * DzVx1 + lam2mu * zty * DzVy1 + lambda * ztz * DzVz1);
                        hAx[TZZ][P] += DBx * Bx * (lambda * ztx * DzVx1 + lambda * zty * DzVy1 + lam2mu * ztz * DzVz1);
                        hAx[TXY][P] += DBx * Bx * (zty * DzVx1 + ztx * DzVy1) * miu;
                        hAx[TXZ][P] += DBx * Bx * (ztz * DzVx1 + ztx * DzVz1) * miu;
                        hAx[TYZ][P] += DBx * Bx * (ztz * DzVy1 + zty * DzVz1) * miu;
                    }
                }
                tempIdx = idy;
                if (tempIdx <= nabs[2] + LFD - 1 || tempIdx >= o->nj + LFD - nabs[3]) {
                    long P = tempIdx < nabs[2] + LFD ? tempIdx - LFD : tempIdx - (o->nj + LFD - nabs[3]) + nabs[2];
                    P = (P * o->NX + idx) * NZ + idz;
                    real *const *Ay = mA[1];
                    real **hAy = o->hA[1];
                    for (int c = 0; c < NF; c++) h[c] -= Ay[c][P] / By;
                    hAy[VX][P] = DetVx * DBy - APDy * Ay[VX][P];
                    hAy[VY][P] = DetVy * DBy - APDy * Ay[VY][P];
                    hAy[VZ][P] = DetVz * DBy - APDy * Ay[VZ][P];
                    hAy[TXX][P] = DetTxx * DBy - APDy * Ay[TXX][P];
                    hAy[TYY][P] = DetTyy * DBy - APDy * Ay[TYY][P];
                    hAy[TZZ][P] = DetTzz * DBy - APDy * Ay[TZZ][P];
#if PLANT(12) /* y auxiliary rate of Txy scaled by the x profile */
                    hAy[TXY][P] = DetTxy * DBx - APDy * Ay[TXY][P];
#else
                    hAy[TXY][P] = DetTxy * DBy - APDy * Ay[TXY][P];
#endif
                    hAy[TXZ][P] = DetTxz * DBy - APDy * Ay[TXZ][P];
                    hAy[TYZ][P] = DetTyz * DBy - APDy * Ay[TYZ][P];
                    if (surf) {
                        hAy[TXX][P] += DBy * By * (lam2mu * ztx * DzVx2 + lambda * zty * DzVy2 + lambda * ztz * DzVz2);
                        hAy[TYY][P] += DBy * By * (lambda * ztx * DzVx2 + lam2mu * zty * DzVy2 + lambda * ztz * DzVz2);
                        hAy[TZZ][P] += DBy * By * (lambda * ztx * DzVx2 + lambda * zty * DzVy2 + lam2mu * ztz * DzVz2);
                        hAy[TXY][P] += DBy * By * (zty * DzVx2 + ztx * DzVy2) * miu;
                        hAy[TXZ][P] += DBy * By * (ztz * DzVx2 + ztx * DzVz2) * miu;
                        hAy[TYZ][P] += DBy * By * (ztz * DzVy2 + zty * DzVz2) * miu;
                    }
                }
                if (idz <= nabs[4] + LFD - 1 || idz >= nk + LFD - nabs[5]) {
                    long P = idz < nabs[4] + LFD ? idz - LFD : idz - (nk + LFD - nabs[5]) + nabs[4];
                    P = (P * o->NX + idx) * NY + idy;
                    real *const *Az = mA[2];
                    real **hAz = o->hA[2];
                    for (int c = 0; c < NF; c++) h[c] -= Az[c][P] / Bz;
                    hAz[VX][P] = DztVx * DBz - APDz * Az[VX][P];
                    hAz[VY][P] = DztVy * DBz - APDz * Az[VY][P];
                    hAz[VZ][P] = DztVz * DBz - APDz * Az[VZ][P];
                    hAz[TXX][P] = DztTxx * DBz - APDz * Az[TXX][P];
                    hAz[TYY][P] = DztTyy * DBz - APDz * Az[TYY][P];
                    hAz[TZZ][P] = DztTzz * DBz - APDz * Az[TZZ][P];
                    hAz[TXY][P] = DztTxy * DBz - APDz * Az[TXY][P];
                    hAz[TXZ][P] = DztTxz * DBz - APDz * Az[TXZ][P];
                    hAz[TYZ][P] = DztTyz * DBz - APDz * Az[TYZ][P];
                }
                for (int c = 0; c < NF; c++) o->hW[c][G] = h[c];
            }
}

/* ---- CalTIMG, calculate.cu:5428-5666 ---------------------------------------
 * Traction image for the top three rows: rebuild the momentum rates of
 * Vx, Vy, Vz from 7-point vectors J*(grad(xi|eta|zeta) . T_row), with the
 * zeta traction forced to zero at the surface and odd images above it. */
static void timg(orc_t *o, real *const *mW, real *const mA[3][NF], int Xvec, int Yvec, int Zvec)
{
    const int NX = o->NX, NY = o->NY, NZ = o->NZ, nk = o->nk;
    const real steph = o->dh;
    const real xstep = steph * Xvec, ystep = steph * Yvec, zstep = steph * Zvec;
    const long xs = (long)NY * NZ, ys = NZ;
    const int *nabs = o->nabs;
    real *const *D = o->drv;
    /* rows of the three momentum equations: (Txx,Txy,Txz) (Txy,Tyy,Tyz) (Txz,Tyz,Tzz) */
    static const int trow[3][3] = {{TXX, TXY, TXZ}, {TXY, TYY, TYZ}, {TXZ, TYZ, TZZ}};
    (void)NX;
#pragma omp parallel for schedule(static)
    for (int idx = LFD; idx < o->ni + LFD; idx++)
        for (int idy = LFD; idy < o->nj + LFD; idy++)
            for (int countZ = 0; countZ < LFD; countZ++) {
                int idz = nk + countZ;
                long G = ((long)idx * NY + idy) * NZ + idz;
                real T3Src = 0.0;
                real rhojac = 1.0 / o->rho[G] / D[JAC][G];
                real APDx = o->APD[0][idx], APDy = o->APD[1][idy];
                real Bx = o->Bt[0][idx], By = o->Bt[1][idy];
                real DBx = o->DB[0][idx], DBy = o->DB[1][idy];
                for (int m = 0; m < 3; m++) {
                    const real *T0 = mW[trow[m][0]], *T1 = mW[trow[m][1]], *T2 = mW[trow[m][2]];
                    real VecTx[7], VecTy[7], VecTz[7];
                    for (int i = 0; i < 2 * LFD + 1; i++) {
                        long r = G - LFD * xs + i * xs;
                        if (o->simplecar) /* CalTIMGmix (calculate.cu:5729-5740) */
                            VecTx[i] = D[JAC][r] * T0[r];
                        else
                            VecTx[i] = D[JAC][r] * (D[XIX][r] * T0[r] + D[XIY][r] * T1[r] + D[XIZ][r] * T2[r]);
                        r = G - LFD * ys + i * ys;
                        if (o->simplecar)
#if PLANT(19) /* the x-row stress in place of the y-row one in CalTIMGmix's eta vector */
                            VecTy[i] = D[JAC][r] * T0[r];
#else
                            VecTy[i] = D[JAC][r] * T1[r];
#endif
                        else
                            VecTy[i] = D[JAC][r] * (D[ETX][r] * T0[r] + D[ETY][r] * T1[r] + D[ETZ][r] * T2[r]);
                        r = G - LFD + i;
#if PLANT(6) /* Jacobian dropped from the zeta traction vector */
                        VecTz[i] = (D[ZTX][r] * T0[r] + D[ZTY][r] * T1[r] + D[ZTZ][r] * T2[r]);
#else
                        VecTz[i] = D[JAC][r] * (D[ZTX][r] * T0[r] + D[ZTY][r] * T1[r] + D[ZTZ][r] * T2[r]);
#endif
                    }
                    for (int j = 1; j <= LFD - (2 - countZ); j++)
#if PLANT(5) /* even image above the surface instead of the odd one */
                        VecTz[LFD + (2 - countZ) + j] = 2.0 * T3Src + VecTz[LFD + (2 - countZ) - j];
#else
                        VecTz[LFD + (2 - countZ) + j] = 2.0 * T3Src - VecTz[LFD + (2 - countZ) - j];
#endif
                    VecTz[LFD + (2 - countZ)] = T3Src;

                    real DxTx = rhojac * drp_fwd(VecTx, LFD, Xvec, xstep);
                    real DyTy = rhojac * drp_fwd(VecTy, LFD, Yvec, ystep);
                    real DzTz = rhojac * drp_fwd(VecTz, LFD, Zvec, zstep);
                    real hv = DxTx / Bx + DyTy / By + DzTz;

                    int tempIdx = idx;
                    if (tempIdx <= nabs[0] + LFD - 1 || tempIdx >= o->ni + LFD - nabs[1]) {
                        long P = tempIdx < nabs[0] + LFD ? tempIdx - LFD : tempIdx - (o->ni + LFD - nabs[1]) + nabs[0];
                        P = (P * NY + idy) * NZ + idz;
                        hv -= mA[0][m][P] / Bx;
                        o->hA[0][m][P] = DBx * DxTx - APDx * mA[0][m][P];
                    }
                    tempIdx = idy;
                    if (tempIdx <= nabs[2] + LFD - 1 || tempIdx >= o->nj + LFD - nabs[3]) {
                        long P = tempIdx < nabs[2] + LFD ? tempIdx - LFD : tempIdx - (o->nj + LFD - nabs[3]) + nabs[2];
                        P = (P * o->NX + idx) * NZ + idz;
                        hv -= mA[1][m][P] / By;
                        o->hA[1][m][P] = DBy * DyTy - APDy * mA[1][m][P];
                    }
                    o->hW[m][G] = hv;
                }
            }
}

#ifdef ORC_DOUBLE
#define RFLOOR floor
#define RFABS fabs
#else
#define RFLOOR floorf
#define RFABS fabsf
#endif

/* ExtractValue (calculate.cu:2586-2613): linear interpolation of one focal
 * series sampled every DT from Tstart, zero outside [Tstart, Tend]; all in
 * `real`.  `avail` samples may be read from stf (the flat series of every
 * later focal point follows, as in the reference's H_Rmnt arrays). */
static real extract_value(const real *stf, long avail, real Tstart, real Tend, real DT, real time)
{
    int PL, PR;
    real value = 0;
    if (time > Tend || time < Tstart)
        value = 0;
    else {
        if (RFABS(time - DT * RFLOOR(time / DT)) < 1e-5) {
            PL = (int)(time / DT);
            PR = PL;
        } else {
            PL = (int)(time / DT);
            PR = PL + 1;
        }
        real vl = PL < avail ? stf[PL] : 0, vr = PR < avail ? stf[PR] : 0;
        value = vl + (vr - vl) * ((time - DT * PL) / DT);
    }
    return value;
}

/* ---- LoadForce (calculate.cu:6130-6191), LoadMoment (6193-6255) and
 * LoadRmom (6257-6328) with InterpFocus (2615-2676) ------------------------ */
static void sources(orc_t *o, int Tindex)
{
    const int NY = o->NY, NZ = o->NZ;
    const real steph = o->dh;
    for (int mid = 0; mid < o->nfrc; mid++) {
        int idx = o->floc[3 * mid], idy = o->floc[3 * mid + 1], idz = o->floc[3 * mid + 2];
        long G = ((long)idx * NY + idy) * NZ + idz;
        real stf = o->fstf[(long)mid * o->nstf_f + Tindex];
        real d = 1.0;
        real A = steph * steph * steph * o->drv[JAC][G];
        A = 1 / (A * o->rho[G]) * d;
#if !PLANT(8) /* PLANT 8: no doubling of a force on the surface row */
        if (idz == o->nk + LFD - 1) A = A * 2.0;
#endif
        o->hW[VX][G] += stf * o->fvec[3 * mid + 0] * A;
        o->hW[VY][G] += stf * o->fvec[3 * mid + 1] * A;
        o->hW[VZ][G] += stf * o->fvec[3 * mid + 2] * A;
    }
    for (int mid = 0; mid < o->nmnt; mid++) {
        int idx = o->mloc[3 * mid], idy = o->mloc[3 * mid + 1], idz = o->mloc[3 * mid + 2];
        long G = ((long)idx * NY + idy) * NZ + idz;
        real stf = o->mstf[(long)mid * o->nstf_m + Tindex];
        real A = 1.0;
#if PLANT(7) /* Jacobian dropped from the moment source */
        A = stf * A / (steph * steph * steph);
#else
        A = stf * A / (steph * steph * steph * o->drv[JAC][G]);
#endif
        o->hW[TXX][G] -= o->mvec[6 * mid + 0] * A;
        o->hW[TYY][G] -= o->mvec[6 * mid + 1] * A;
        o->hW[TZZ][G] -= o->mvec[6 * mid + 2] * A;
        o->hW[TXY][G] -= o->mvec[6 * mid + 3] * A;
        o->hW[TXZ][G] -= o->mvec[6 * mid + 4] * A;
        o->hW[TYZ][G] -= o->mvec[6 * mid + 5] * A;
    }
    if (o->nfoc) {
        /* stage time (currT + tinc*0.5)*stept, calculate.cu:1422; Tindex = 2*currT+tinc */
        real time = (Tindex * 0.5) * o->dt;
        real Tstart = 0, Tend = o->fdt * (o->fnt - 1);
        const long all = (long)o->nfoc * o->fnt;
        for (int mid = 0; mid < o->nfoc; mid++) {
            int idx = o->rmloc[3 * mid], idy = o->rmloc[3 * mid + 1], idz = o->rmloc[3 * mid + 2];
            long G = ((long)idx * NY + idy) * NZ + idz;
            long src = (long)mid * o->fnt;
            real m[6];
            for (int c = 0; c < 6; c++)
                m[c] = extract_value(o->rmser + c * all + src, all - src, Tstart, Tend, o->fdt, time);
            real A = 1.0;
            A = A / (steph * steph * steph * o->drv[JAC][G]);
            o->hW[TXX][G] -= m[0] * A;
            o->hW[TYY][G] -= m[1] * A;
            o->hW[TZZ][G] -= m[2] * A;
            o->hW[TXY][G] -= m[3] * A;
            o->hW[TXZ][G] -= m[4] * A;
            o->hW[TYZ][G] -= m[5] * A;
        }
    }
}

/* ---- Iteration{Begin,Inner,Final,FinalPV}, calculate.cu:6330-6852 -------- */
static void update(orc_t *o, int k)
{
    const int NY = o->NY, NZ = o->NZ;
    real alpha = k < 3 ? RK4A[k] : 0, beta = RK4B[k];
    alpha *= o->dt;
    beta *= o->dt;
#pragma omp parallel for schedule(static)
    for (int idx = LFD; idx < o->ni + LFD; idx++)
        for (int idy = LFD; idy < o->nj + LFD; idy++)
            for (int idz = LFD; idz < o->nk + LFD; idz++) {
                long G = ((long)idx * NY + idy) * NZ + idz;
                for (int c = 0; c < NF; c++) {
                    real h = o->hW[c][G];
                    if (k == 0) {
                        o->W[c][G] = o->FW[c][G] + alpha * h;
                        o->tW[c][G] = o->FW[c][G] + beta * h;
                    } else if (k < 3) {
                        o->W[c][G] = o->FW[c][G] + alpha * h;
                        o->tW[c][G] = o->tW[c][G] + beta * h;
                    } else {
                        o->W[c][G] = o->tW[c][G] + beta * h;
                    }
                }
                if (k == 3 && o->pgv_on && idz == o->nk + LFD - 1) {
                    long xi = (long)idx * NY + idy;
                    for (int c = 0; c < 3; c++)
                        o->pgv[c][xi] = OMAX(OABS(o->W[c][G]), OABS(o->pgv[c][xi]));
                }
            }
    /* auxiliary fields over their packed ranges */
    for (int a = 0; a < 3; a++) {
        long n = o->asz[a];
        if (!n) continue;
        /* only entries at valid (y,z) / (x,z) / (x,y) positions are ever
         * nonzero; updating the ghost-position entries too is harmless
         * (0 + coef*0 = 0) and keeps the loop flat */
        for (int c = 0; c < NF; c++) {
            real *FA = o->FA[a][c], *Aa = o->A[a][c], *tA = o->tA[a][c], *hA = o->hA[a][c];
#pragma omp parallel for schedule(static)
            for (long p = 0; p < n; p++) {
                if (k == 0) {
                    Aa[p] = FA[p] + alpha * hA[p];
                    tA[p] = FA[p] + beta * hA[p];
                } else if (k < 3) {
                    Aa[p] = FA[p] + alpha * hA[p];
                    tA[p] = tA[p] + beta * hA[p];
                } else {
                    Aa[p] = tA[p] + beta * hA[p];
                }
            }
        }
    }
}

static void record(orc_t *o)
{
    const int NY = o->NY, NZ = o->NZ;
    if (o->currT >= o->nt) return;
    for (int r = 0; r < o->nrecv; r++) {
        long G = ((long)o->rloc[3 * r] * NY + o->rloc[3 * r + 1]) * NZ + o->rloc[3 * r + 2];
        for (int c = 0; c < NF; c++)
            o->seis[((long)c * o->nrecv + r) * o->nt + o->currT] = o->FW[c][G];
    }
}

/* F/B signs per (step % 4, stage), simulation.cc:264-641: x and y share a
 * sign; x/y alternate with the stage parity and flip for steps 2,3 mod 4; z
 * alternates with the stage parity and flips for steps 1,2 mod 4. */
static void fb_signs(int currT, int k, int *sx, int *sz)
{
    int p = currT % 4;
    *sx = (((p >= 2) ? 1 : 0) ^ (k & 1)) ? -1 : 1;
#if PLANT(9) /* z signs follow the x pattern (wrong in step phases 1 and 3) */
    *sz = *sx;
#else
    *sz = (((p == 1 || p == 2) ? 1 : 0) ^ (k & 1)) ? -1 : 1;
#endif
}

static void stage(orc_t *o)
{
    int k = o->stage, sx, sz;
    fb_signs(o->currT, k, &sx, &sz);
    int tinc = k == 0 ? 0 : (k < 3 ? 1 : 2);
    int Tindex = 2 * o->currT + tinc; /* calculate.cu:1404-1421 */
    real *const *mW = k == 0 ? o->FW : o->W;
    real *mA[3][NF];
    for (int a = 0; a < 3; a++)
        for (int c = 0; c < NF; c++) mA[a][c] = k == 0 ? o->FA[a][c] : o->A[a][c];
    rhs(o, mW, (real *const(*)[NF])mA, sx, sx, sz);
    timg(o, mW, (real *const(*)[NF])mA, sx, sx, sz);
    sources(o, Tindex);
    update(o, k);
    if (k == 3) {
        /* wavesyn(FW,W) + abssyn(3), simulation.cc:310-311 */
        for (int c = 0; c < NF; c++) memcpy(o->FW[c], o->W[c], o->N * sizeof(real));
        for (int a = 0; a < 3; a++)
            for (int c = 0; c < NF; c++)
                if (o->asz[a]) memcpy(o->FA[a][c], o->A[a][c], o->asz[a] * sizeof(real));
        record(o);
        o->currT++;
        o->stage = 0;
    } else {
        o->stage = k + 1;
    }
}

int orc_run_stages(orc_t *o, int nstages)
{
    for (int s = 0; s < nstages; s++) stage(o);
    return 0;
}

int orc_run(orc_t *o, int nsteps)
{
    if (o->stage != 0) return -1;
    return orc_run_stages(o, 4 * nsteps);
}

int orc_current_step(const orc_t *o) { return o->currT; }
int orc_current_stage(const orc_t *o) { return o->stage; }

static real *const *field_set(const orc_t *o, int which)
{
    switch (which) {
    case 0: return o->FW;
    case 1: return o->W;
    case 2: return o->tW;
    case 3: return o->hW;
    default: return NULL;
    }
}

int orc_get_field(const orc_t *o, int which, int comp, real *out)
{
    real *const *s = field_set(o, which);
    if (!s || comp < 0 || comp >= NF) return -1;
    memcpy(out, s[comp], o->N * sizeof(real));
    return 0;
}

int orc_set_field(orc_t *o, int which, int comp, const real *in)
{
    real *const *s = field_set(o, which);
    if (!s || comp < 0 || comp >= NF) return -1;
    memcpy(s[comp], in, o->N * sizeof(real));
    return 0;
}

long orc_aux_size(const orc_t *o, int axis) { return (axis >= 0 && axis < 3) ? o->asz[axis] : -1; }

static real *aux_ptr(const orc_t *o, int which, int axis, int comp)
{
    if (axis < 0 || axis > 2 || comp < 0 || comp >= NF) return NULL;
    switch (which) {
    case 0: return o->FA[axis][comp];
    case 1: return o->A[axis][comp];
    case 2: return o->tA[axis][comp];
    case 3: return o->hA[axis][comp];
    default: return NULL;
    }
}

int orc_get_aux(const orc_t *o, int which, int axis, int comp, real *out)
{
    real *p = aux_ptr(o, which, axis, comp);
    if (!p) return o->asz[axis >= 0 && axis < 3 ? axis : 0] == 0 ? 0 : -1;
    memcpy(out, p, o->asz[axis] * sizeof(real));
    return 0;
}

int orc_set_aux(orc_t *o, int which, int axis, int comp, const real *in)
{
    real *p = aux_ptr(o, which, axis, comp);
    if (!p) return o->asz[axis >= 0 && axis < 3 ? axis : 0] == 0 ? 0 : -1;
    memcpy(p, in, o->asz[axis] * sizeof(real));
    return 0;
}

int orc_get_seismo(const orc_t *o, real *out)
{
    if (o->nrecv) memcpy(out, o->seis, (size_t)9 * o->nrecv * o->nt * sizeof(real));
    return 0;
}

int orc_get_pgv(const orc_t *o, real *out)
{
    size_t n = (size_t)o->NX * o->NY;
    for (int c = 0; c < 3; c++) memcpy(out + c * n, o->pgv[c], n * sizeof(real));
    return 0;
}

int orc_get_cov(const orc_t *o, real *covx, real *covy)
{
    size_t n = (size_t)o->NX * o->NY * 9;
    memcpy(covx, o->covx, n * sizeof(real));
    memcpy(covy, o->covy, n * sizeof(real));
    return 0;
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
