/*
 * cgfd3d.h -- C-ABI of the MI355X (gfx950) time-stepping engine for the
 * CGFD3D curvilinear-grid elastic-wave solver.
 *
 * The reference (fes478/CoPara-CGFD3D) has no plugin or FFI layer: its hot
 * path sits behind the C++ class ChildProcs (src/typedata.h:114-217), which
 * owns the device buffers, receives the host-prepared parameter arrays
 * (ParH2D, src/calculate.cu:2050-2228), computes the free-surface closure
 * (VelCoeff, calculate.cu:1336-1353), advances one RK stage per call (RKite,
 * calculate.cu:1355-1599) under the forward/backward schedule of
 * src/simulation.cc:252-717, and hands receiver traces and peak ground
 * velocity back to the host (PWpick calculate.cu:2780-2828, SynPV 1709-1743).
 * Each entry point below names the member it replaces.
 *
 * Conventions
 *  - All functions return 0 on success and a negative CGFD_E* code on
 *    failure; cgfd_last_error() gives the message.  Nothing aborts.
 *  - Host arrays are borrowed for the duration of the call only.
 *  - Array element type is `float` for a CGFD_FP32 context and `double` for a
 *    CGFD_FP64 context (the reference's Real, typenew.h:13-30); `void*`
 *    arguments below point to that type.
 *  - 3-D arrays use the reference layout: flat [x][y][z], z fastest, with 3
 *    ghost cells on each side: NX=ni+6, NY=nj+6, NZ=nk+6, valid indices
 *    [3, n+3); the free surface is the last valid z row (z = nk+2).
 *  - Point locations (sources, receivers) are ghost-inclusive (x,y,z)
 *    triples, as the reference's locx/locy/locz (calculate.cu:6146-6148).
 *  - Wavefield component order: Vx Vy Vz Txx Tyy Tzz Txy Txz Tyz
 *    (typenew.h:265-270).
 *  - One context drives one GPU from one host thread on one HIP stream.
 */
#ifndef CGFD3D_H
#define CGFD3D_H

#ifdef __cplusplus
extern "C" {
#endif

#define CGFD_FP32 0
#define CGFD_FP64 1

#define CGFD_OK 0
#define CGFD_EINVAL (-1)    /* bad argument or inconsistent configuration */
#define CGFD_ESTATE (-2)    /* call out of order (e.g. run before init) */
#define CGFD_EHIP (-3)      /* HIP runtime error */
#define CGFD_ENOMEM (-4)    /* device allocation failed */
#define CGFD_ERANGE (-5)    /* point outside the valid domain */

typedef struct cgfd_ctx cgfd_ctx;

/* Grid and time stepping (cindx typenew.h:135-141, steph/stept/nt of the
 * ChildProcs constructor calculate.cu:107-111). */
typedef struct cgfd_dims {
    int ni, nj, nk; /* valid cells along x, y, z */
    double dh;      /* grid step (steph) */
    double dt;      /* time step (stept) */
    int nt;         /* number of time steps the run is sized for */
} cgfd_dims;

/* ChildProcs::ChildProcs (calculate.cu:107-1003): allocate device state on
 * `device` (HIP ordinal).  precision: CGFD_FP32 or CGFD_FP64.  Returns
 * CGFD_EINVAL (before touching the GPU) when one padded field, about
 * (ni+67)*(nj+3)*(nk+8) elements, would exceed 4 GiB: the stage kernels
 * address a field with 32-bit byte offsets. */
int cgfd_create(const cgfd_dims *dims, int precision, int device, cgfd_ctx **out);
/* ChildProcs::~ChildProcs (calculate.cu:1016-1334) */
void cgfd_destroy(cgfd_ctx *ctx);
const char *cgfd_last_error(const cgfd_ctx *ctx);

/* ParH2D for derivF (calculate.cu:2050-2228): 10 arrays xix xiy xiz etax
 * etay etaz zetax zetay zetaz jac, NX*NY*NZ each, reference layout. */
int cgfd_set_metrics(cgfd_ctx *ctx, const void *const *drv);
/* ParH2D for mdparF: P velocity alpha, S velocity beta, density rho. */
int cgfd_set_media(cgfd_ctx *ctx, const void *alpha, const void *beta, const void *rho);
/* ParH2D for apara (typenew.h:277-282): nabs[6] = layers on x1 x2 y1 y2 z1 z2
 * (z2 = top, forced to 0 under the free surface, absorb.cpp:142-148);
 * tabs[9] = APDx APDy APDz Bx By Bz DBx DBy DBz with lengths NX NY NZ. */
int cgfd_set_pml(cgfd_ctx *ctx, const int *nabs, const void *const *tabs);
/* ParH2D for forceF (typedata.h:70-77): n point forces; locs[3n];
 * vec[3n] = fx fy fz per source (already scaled by the force factor);
 * stf[n*nstf] sampled every dt/2 (source.cpp:408-416), nstf >= 2*nt+1. */
int cgfd_set_forces(cgfd_ctx *ctx, int n, const int *locs, const void *vec, const void *stf, int nstf);
/* ParH2D for momentF (typedata.h:79-86): vec[6n] = mxx myy mzz mxy mxz myz. */
int cgfd_set_moments(cgfd_ctx *ctx, int n, const int *locs, const void *vec, const void *stf, int nstf);
/* ParH2D for RmomF (typedata.h:88-97) with the host series of
 * source_type 1 (source.cpp:43-58): n focal points; series[6][n][fnt] =
 * mxx myy mzz mxy mxz myz moment-rate samples every fdt seconds from t = 0
 * (focal_import layout, seisnc.cpp:1148).  Each stage applies their linear
 * interpolation at the stage time (InterpFocus calculate.cu:2615-2676,
 * ExtractValue 2586-2613) as LoadRmom does (6257-6328): T -= m(t)/(h^3 J),
 * after the moment sources. */
int cgfd_set_focal(cgfd_ctx *ctx, int n, const int *locs, const void *series, int fnt, double fdt);
/* C2DPointPick (calculate.cu:2678-2746): receivers recording all 9
 * components of the step-end wavefield every step. */
int cgfd_set_receivers(cgfd_ctx *ctx, int n, const int *locs);
/* C2DSnapPick (calculate.cu:2230-2358) for the snapshot boxes of
 * seisplot.cpp:350-361: nsnap boxes, spec[11*nsnap] = xs ys zs (1-based
 * valid-grid start), xn yn zn (point counts), xi yi zi (index intervals),
 * tinv, cmp (1 = Vx Vy Vz, 2 = the six stresses, 3 = all nine).  Points sit
 * at ghost-inclusive (xs+2+k*xi, ys+2+m*yi, zs+2+n*zi) and must lie in the
 * NX*NY*NZ array.  Replaces any earlier boxes. */
int cgfd_set_snapshots(cgfd_ctx *ctx, int nsnap, const int *spec);
/* npoints = xn*yn*zn and ntime = ceil(nt/tinv) of snapshot isnap. */
int cgfd_snap_shape(const cgfd_ctx *ctx, int isnap, int *npoints, int *ntime);
/* SWpick (calculate.cu:2360-2444) + SWgather: after every step t with
 * t % tinv == 0 the step-end field FW is recorded into time slot t/tinv.
 * Returns component comp (0..8, within the group cmp selects) as
 * [npoints][ntime], point index k*yn*zn + m*zn + n (seisplot.cpp:513-518). */
int cgfd_get_snap(cgfd_ctx *ctx, int isnap, int comp, void *out);
/* output_peakvel (PVF, calculate.cu:134): running max |V| on the surface. */
int cgfd_set_pgv(cgfd_ctx *ctx, int on);
/* hybrid_grid = 1: the ChildProcs members HyGrid/ConIndex (typedata.h:143),
 * which select CalDiff/CalWave instead of CalDiffCL/CalWaveCL
 * (calculate.cu:1469-1505).  conindex is the reference's ConIndex, a z index
 * counted with the 3 ghost planes (gridmesh.cpp:346): below it the rates use
 * CalWave's Cartesian contraction of xi_x, eta_y and zeta_z
 * (calculate.cu:3910-3956), from it up the curvilinear one.  The default,
 * 3 (= nk1), is the all-curvilinear grid of hybrid_grid = 0.  CGFD_EINVAL
 * unless 3 <= conindex <= nk (the three surface rows stay curvilinear). */
int cgfd_set_hybrid(cgfd_ctx *ctx, int conindex);
/* The reference's SimpleCar build (its Makefile default), selected there at
 * compile time: CalDiffmix/CalWavemix/CalTIMGmix in place of
 * CalDiff/CalWave/CalTIMG and VelPDcoeff with e11 = e22 = 1 and the other
 * xi/eta entries 0 (calculate.cu:1471-1480, 1509-1511, 3362-3375): the
 * xi/eta metrics are taken as the identity whatever the metric arrays hold.
 * Call before cgfd_set_metrics; on grids whose xi/eta metrics are the
 * identity (uniform x/y spacing) both builds give the same values. */
int cgfd_set_simplecar(cgfd_ctx *ctx, int on);
/* No reference counterpart: ctx uses src's metrics and media (device arrays)
 * instead of a copy of its own, for many scenarios on one medium (the
 * surrogate's training runs).  Both must be on one device, of one precision
 * and grid size; src must have its metrics and media, ctx none yet.  The
 * arrays are freed with the last context that holds them; while more than
 * one does, none of them may set metrics or media.  In a fused batch, such
 * grids of one tile width and z-chunk length run their copies of each tile
 * together, so the shared lines are read from cache. */
int cgfd_share_medium(cgfd_ctx *ctx, const cgfd_ctx *src);

/* VelCoeff (calculate.cu:1336-1353) plus state reset: zero the wavefield,
 * auxiliary fields, traces and PGV; compute the traction-free closure
 * matrices; set the step counter to 0. */
int cgfd_init(cgfd_ctx *ctx);

/* The time loop of simulation.cc:252-717 for one process and one GPU:
 * advance nsteps whole steps (4 RK stages each, F/B pattern chosen by
 * step % 4) from the current step.  Asynchronous on the context's stream. */
int cgfd_run(cgfd_ctx *ctx, int nsteps);
/* One RKite call (calculate.cu:1355-1599) per stage; the 4th stage of a
 * step finishes it (state copy, traces, step counter). */
int cgfd_run_stages(cgfd_ctx *ctx, int nstages);
/* Batch of independent simulations (no reference counterpart; the
 * reference runs one scenario per job): advance each of the n contexts by
 * nsteps steps.  A fusable list (cgfd_batch_fusable) runs as one launch per
 * kernel and stage for all its grids; otherwise stage k of every context is
 * enqueued on its own stream before stage k+1 of any, so small grids share
 * the GPU.  Results equal those of running each context alone. */
int cgfd_run_batch(cgfd_ctx *const *ctxs, int n, int nsteps);
/* No counterpart in the reference: cgfd_run_batch's fused path for callers
 * that cannot allocate device memory (code under graph capture, serving
 * loops, harnesses that forbid hipMalloc).  The list must be one that
 * cgfd_batch_fusable accepts, of at most 64 grids; the batch's argument
 * images go into a fixed array of the library (513 KiB per device, held
 * until the process ends) instead of a buffer it allocates, and nothing is
 * allocated.  Results equal cgfd_run_batch's.  Every call on a device shares
 * that device's array: a call waits for the previous call's launches before
 * it rewrites it, but calls that overlap on one device (from several host
 * threads, or while another call is being set up) race.  CGFD_EINVAL for a
 * longer or unfusable list, with the reason in cgfd_last_error of the first
 * context. */
int cgfd_run_batch_static(cgfd_ctx *const *ctxs, int n, int nsteps);
/* 1 if cgfd_run_batch runs this list as one launch per kernel (no reference
 * counterpart): 2 to 4096 distinct contexts on one device with the same
 * precision, the default stage variant, the same metric mode and the same
 * step phase (step % 4), none timing its kernels.  Each grid runs at its
 * own tile width (cgfd_tile_width): one stage launch per width.  Grid sizes,
 * media, PML, sources, receivers and snapshots may differ.  Such a batch
 * runs on the first context's stream, which waits for the others' queued
 * work; a grid keeps the z-chunk count cgfd_autotune or cgfd_set_tuning gave
 * it, the others get the batch's own (about 2048 workgroups per launch); the
 * other streams wait for the batch.  Otherwise 0 (or CGFD_EINVAL for a null
 * list). */
int cgfd_batch_fusable(cgfd_ctx *const *ctxs, int n);
/* Autotuner (no reference counterpart): time trial_steps steps for a few
 * z-chunk counts around the default and keep the fastest (*chosen); in FP32
 * and unless cgfd_set_tile_width fixed one, then the other tile width of the
 * default stage kernel at that count, kept if faster; with trial_steps >= 4
 * and nt >= 8, also
 * time 4 eager steps against 4 replayed ones and switch graph replay
 * (cgfd_set_graph) on if it is faster.  Then re-initialise the context.
 * Call right after cgfd_init. */
int cgfd_autotune(cgfd_ctx *ctx, int trial_steps, int *chosen);
/* Graph replay (no reference counterpart; for small grids whose steps are
 * bound by the gaps between dependent launches): with on != 0, cgfd_run
 * captures the launches of 4 whole steps once per starting phase (step % 4
 * fixes the forward/backward pattern) into a HIP graph and replays it; the
 * step number reaches the kernels through a device-side counter, so results
 * equal eager stepping.  Steps left over (fewer than 4) run eagerly.
 * Inactive while profiling. */
int cgfd_set_graph(cgfd_ctx *ctx, int on);
int cgfd_sync(cgfd_ctx *ctx);
int cgfd_current_step(const cgfd_ctx *ctx);
int cgfd_current_stage(const cgfd_ctx *ctx);

/* State access in the reference layout.  which: 0 = FW (step-start field),
 * 1 = W (output of the last stage), 2 = tW (RK accumulator).  Ghost cells
 * read back as 0 and are ignored on write. */
int cgfd_get_field(cgfd_ctx *ctx, int which, int comp, void *out);
int cgfd_set_field(cgfd_ctx *ctx, int which, int comp, const void *in);
/* ADE auxiliary fields in the reference's packed layout
 * (calculate.cu:325-327): axis 0 [(nabs0+nabs1)][NY][NZ],
 * axis 1 [(nabs2+nabs3)][NX][NZ], axis 2 [(nabs4+nabs5)][NX][NY];
 * which as above (0 = FA, 1 = A, 2 = tA). */
long cgfd_aux_size(const cgfd_ctx *ctx, int axis);
int cgfd_get_aux(cgfd_ctx *ctx, int which, int axis, int comp, void *out);
int cgfd_set_aux(cgfd_ctx *ctx, int which, int axis, int comp, const void *in);

/* PWpick results (calculate.cu:2780-2828): [9][nrecv][nt]. */
int cgfd_get_seismo(cgfd_ctx *ctx, void *out);
/* SynPV (calculate.cu:1709-1743): [3][NX][NY] (Vx, Vy, Vz maxima). */
int cgfd_get_pgv(cgfd_ctx *ctx, void *out);
/* matVx2Vz / matVy2Vz of VelPDcoeff (calculate.cu:3320-3428): [NX][NY][9]. */
int cgfd_get_cov(cgfd_ctx *ctx, void *covx, void *covy);

/* ---- NetCDF output (seisnc.cpp) -------------------------------------------
 * Files carry the reference's dimensions, variables and global attributes
 * and are written in the netCDF classic format with 64-bit offsets, which
 * netCDF readers open like the reference's NETCDF4 files.  Real is float
 * (CGFD_FP32) or double (CGFD_FP64). */
/* point_export_def + point_export (seisnc.cpp:1229-1365): seismo.nc with
 * Vx..Tyz[TimeNum][PointNum], currTime[TimeNum], Pnum, Lnum (int), posx,
 * posy, posz [PointNum]; wave[9] point to [nrec][npnt] arrays. */
int cgfd_nc_seismo(const char *path, int precision, int nrec, int npnt, const void *const *wave,
                   const void *currtime, const int *pnum, const int *lnum, const void *posx, const void *posy,
                   const void *posz);
/* snap_export_def + snap_export (seisnc.cpp:1275-1490): snap_NNN.nc with
 * the components cmp selects as [TimeNum][i][j][k], currTime, and the
 * IdxStart/IdxNumber/IdxInterval/TimeInterval/Content attributes; spec is
 * the 11-integer snapshot box of cgfd_set_snapshots; wave[c] (unused
 * components may be NULL) point to [nrec][xn][yn][zn] arrays. */
int cgfd_nc_snap(const char *path, int precision, const int *spec, int nrec, const void *const *wave,
                 const void *currtime);
/* PV_export (seisnc.cpp:1671-1712): peakvel.nc with Vx, Vy, Vz [i][j]. */
int cgfd_nc_peakvel(const char *path, int precision, int ni, int nj, const void *vx, const void *vy, const void *vz);

/* End-of-run outputs of a context, as the reference's master writes them
 * (simulation.cc:319-360).  seismo.nc: every step `it` goes to record
 * it/tinv (tinv_of_seismo) with currTime = (it+1)*dt, the last step of a
 * record winning (seisplot.cpp:672-684); pnum, lnum, posx/y/z may be NULL
 * (point index, line 1, zero positions).  snap_NNN.nc: record s = time slot
 * s with currTime = (s*tinv+1)*dt (seisplot.cpp:700-760).  peakvel.nc:
 * vz_from_vy = 1 reproduces the reference's PV_export, which fills Vz from
 * the Vy map (seisnc.cpp:1703-1706); 0 writes the true Vz map. */
int cgfd_write_seismo(cgfd_ctx *ctx, const char *path, int tinv, const int *pnum, const int *lnum, const void *posx,
                      const void *posy, const void *posz);
int cgfd_write_snap(cgfd_ctx *ctx, int isnap, const char *path);
int cgfd_write_peakvel(cgfd_ctx *ctx, const char *path, int vz_from_vy);

/* Measurement hooks (no reference counterpart; the reference only prints
 * wall-clock ETA, simulation.cc:710-726).  With profiling on, every stage
 * kernel launch is bracketed by HIP events on the context's stream;
 * cgfd_profile_read returns the summed milliseconds and launch counts of the
 * fused stage kernel and of the traction-image kernel. */
int cgfd_profile(cgfd_ctx *ctx, int on);
int cgfd_profile_read(cgfd_ctx *ctx, double *ms_stage, long *n_stage, double *ms_timg, long *n_timg);
/* Launch shape of the fused stage kernel: z-chunks per column sweep and
 * rows per workgroup (0 keeps the current value). */
int cgfd_set_tuning(cgfd_ctx *ctx, int zchunks, int rows);
/* Tile width of the default (LDS) stage kernel, no reference counterpart:
 * 64 columns x 4 rows per workgroup, or 32 x 8, which leaves fewer lanes idle
 * on rows whose length is 1-32 past a multiple of 64 (80 cells: 96 lanes
 * instead of 128) but costs about 6 % more per lane.  0 (the default) takes
 * 32 when 64-column tiles need over 8 % more lanes; FP64 contexts take 0 or
 * 64.  Results do not depend on it.
 * cgfd_tile_width returns the width the next launch uses. */
int cgfd_set_tile_width(cgfd_ctx *ctx, int width);
int cgfd_tile_width(const cgfd_ctx *ctx);
/* Stage kernel variant: 0 = xy planes staged in LDS (default), 1 = in-plane
 * stencil taps loaded directly through the vector caches.  Both produce
 * identical results. */
int cgfd_set_variant(cgfd_ctx *ctx, int variant);
/* Algorithmic HBM bytes of one RK stage of this configuration (the
 * 192 B/cell + 144 B/PML-cell-axis model, FP32; doubled for FP64; 168 B/cell
 * when the constant-horizontal-metrics kernel runs). */
double cgfd_stage_bytes(const cgfd_ctx *ctx);
/* 1 when set_metrics found xix = etay = 1 and xiy = xiz = etax = etaz = 0
 * (either sign) in every valid cell (fixed x/y spacing, as the reference's
 * vmap meshes have) and the default stage kernel therefore loads only the
 * three zeta metrics.  Results have identical values either way (only the
 * sign of an exactly-zero result can differ).  cgfd_set_specialize(ctx, 0)
 * forces the general kernel. */
int cgfd_metrics_specialized(const cgfd_ctx *ctx);
int cgfd_set_specialize(cgfd_ctx *ctx, int on);
/* The HIP stream (hipStream_t) the context launches on. */
void *cgfd_stream(const cgfd_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
