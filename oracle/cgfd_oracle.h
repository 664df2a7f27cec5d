/*
 * cgfd_oracle.h -- CPU oracle for the CGFD3D time-stepping hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (fes478/CoPara-CGFD3D, src/calculate.cu kernels CalDiffCL,
 * CalWaveCL, CalTIMG, VelPDcoeff, LoadForce, LoadMoment, LoadRmom, Iteration*, driven by
 * the F/B schedule of src/simulation.cc:252-717).  Only tests, the smoke check
 * and the CPU-baseline leg of the benchmark may load it, and only as the
 * checker.  The product (the HIP library) never calls it.
 *
 * Build: compiled twice from the same source, once with `real` = float
 * (liborc32.so, the reference's default `Real`, typenew.h:14-21) and once with
 * `real` = double (liborc64.so, the reference's -DTypeDouble build,
 * typenew.h:22-30).  -ffp-contract=off: no fused multiply-adds, so every
 * floating-point operation rounds exactly as the reference source spells it.
 *
 * Layout (reference layout, typenew.h:33 / calculate.cu:3648): every 3-D array
 * is flat [x][y][z], z fastest, with LFD=3 ghost cells on every side:
 * NX=ni+6, NY=nj+6, NZ=nk+6, valid indices [3, n+3).  The free surface is the
 * last valid z row (z = nk+2).  Wavefield ghosts stay zero.
 *
 * Field order everywhere: Vx Vy Vz Txx Tyy Tzz Txy Txz Tyz (typenew.h:265-270).
 */
#ifndef CGFD_ORACLE_H
#define CGFD_ORACLE_H

#ifdef ORC_DOUBLE
typedef double real;
#else
typedef float real;
#endif

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc orc_t;

orc_t *orc_create(int ni, int nj, int nk, real dh, real dt, int nt);
void orc_destroy(orc_t *o);

/* 10 arrays xix xiy xiz etax etay etaz zetax zetay zetaz jac, NX*NY*NZ each */
int orc_set_metrics(orc_t *o, const real *const *drv);
int orc_set_media(orc_t *o, const real *alpha, const real *beta, const real *rho);
/* tabs: APDx APDy APDz Bx By Bz DBx DBy DBz (lengths NX,NY,NZ,...) */
int orc_set_pml(orc_t *o, const int *nabs, const real *const *tabs);
/* locs: 3*n ghost-inclusive indices (x,y,z); vec: n*3 (fx fy fz) or n*6
 * (mxx myy mzz mxy mxz myz); stf: n*nstf samples at dt/2 */
int orc_set_forces(orc_t *o, int n, const int *locs, const real *vec, const real *stf, int nstf);
int orc_set_moments(orc_t *o, int n, const int *locs, const real *vec, const real *stf, int nstf);
/* focal mechanism (source_type 1): n points; series[6][n][fnt] = mxx myy
 * mzz mxy mxz myz sampled every fdt from t = 0, interpolated at each stage's
 * time (InterpFocus, calculate.cu:2615-2676) */
int orc_set_focal(orc_t *o, int n, const int *locs, const real *series, int fnt, real fdt);
int orc_set_receivers(orc_t *o, int n, const int *locs);
int orc_set_pgv(orc_t *o, int on);
/* hybrid_grid = 1: the reference's ConIndex (z index with the 3 ghosts);
 * below it CalWave's Cartesian contraction applies (calculate.cu:3910-3956).
 * -1 unless 3 <= conindex <= nk: the three surface rows stay curvilinear. */
int orc_set_conindex(orc_t *o, int conindex);
/* the SimpleCar build (xi/eta metrics taken as the identity in CalWavemix,
 * CalTIMGmix and VelPDcoeff, calculate.cu:1471-1480, 3362-3375) */
int orc_set_simplecar(orc_t *o, int on);

/* Zero the state and compute the free-surface closure matrices (VelPDcoeff). */
int orc_init(orc_t *o);

/* Advance by whole steps / by single RK stages (stage 3 completes a step). */
int orc_run(orc_t *o, int nsteps);
int orc_run_stages(orc_t *o, int nstages);
int orc_current_step(const orc_t *o);
int orc_current_stage(const orc_t *o);

/* which: 0=FW (step-start state), 1=W (last stage output), 2=tW (RK
 * accumulator), 3=hW (last stage rate).  Full NX*NY*NZ arrays. */
int orc_get_field(const orc_t *o, int which, int comp, real *out);
int orc_set_field(orc_t *o, int which, int comp, const real *in);
/* ADE auxiliary fields, reference packed layout (calculate.cu:325-327):
 * axis 0 [(nabs0+nabs1)][NY][NZ], axis 1 [(nabs2+nabs3)][NX][NZ],
 * axis 2 [(nabs4+nabs5)][NX][NY]; which as above (0=FA,1=A,2=tA,3=hA). */
long orc_aux_size(const orc_t *o, int axis);
int orc_get_aux(const orc_t *o, int which, int axis, int comp, real *out);
int orc_set_aux(orc_t *o, int which, int axis, int comp, const real *in);

/* seismograms [9][nrecv][nt]; pgv [3][NX][NY]; closure matrices [NX][NY][9] */
int orc_get_seismo(const orc_t *o, real *out);
int orc_get_pgv(const orc_t *o, real *out);
int orc_get_cov(const orc_t *o, real *covx, real *covy);

int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
