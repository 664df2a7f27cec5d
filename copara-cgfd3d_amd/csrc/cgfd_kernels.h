// cgfd_kernels.h -- argument blocks and launchers of the gfx950 kernels.
// Included by the kernel translation unit and by the C-ABI implementation.
#pragma once
#include "cgfd_device.h"

namespace cgfd {

// Arguments of one fused RK stage (RKite, calculate.cu:1355-1599).
// Buffer sets hold the 9 wavefield components at stride g.fs.
template <typename R> struct StageArgs {
    Geom g;
    const R *win; // stage input (FW for stage 0, else the previous stage output)
    const R *fw;  // step-start field FW (read by stages 1 and 2)
    R *wout;      // stage output (FW itself in stage 3)
    R *tw;        // RK accumulator
    const R *met; // 10 metric fields at stride g.fs
    const R *med; // lam2mu, miu, 1/rho, rho at stride g.fs
    const R *pml; // APD B DB for x (ni each), then y (nj each), then z (nk each)
    int nabs[6];
    int npml[3];   // packed PML widths per axis
    long astr[3];  // stride between the 9 components of one auxiliary set
    const R *ain[3];
    const R *af[3];
    R *aout[3];
    R *at[3];
    const R *timg; // traction-image results for the top 3 rows, 9 slots
    long tstr;
    const R *cov;  // free-surface closure matrices, 18 slots of [nj][nxp]
    long cstr;
    const int *col_first; // per (y*ni+x) column: first source entry or -1
    const int4 *sent;     // source entries (z, kind 0 force / 1 moment / 2 focal, source id, column),
                          // sentinel column -1
    const R *scoef;       // 8 coefficients per entry
    const R *fstf;
    const R *mstf;
    const R *rstf; // focal points: interpolated moment rates [point][6][nstf_r]
    int nstf_f, nstf_m, nstf_r;
    R *pgv; // 3 slots of [nj][nxp] or null
    double h, rh;
    R hr;
    R alpha, beta;
    int zlen;
    int hmet; // 1: horizontal metrics are the constants 1/0 (variant 0 loads only the zeta metrics)
    int zcon; // hybrid grid: planes z < zcon take the Cartesian contraction (0: none)
    int tile; // LDS kernel tile width, 64 or 32 columns (host side: picks the kernel)
    const int *step; // device step counter: the current step currT (Tindex = 2*currT + tinc)
};

template <typename R> struct TimgArgs {
    Geom g;
    const R *win;
    const R *met;
    const R *med; // lam2mu, miu, 1/rho, rho
    const R *pml;
    int nabs[6];
    int npml[2];
    long astr[2];
    const R *ain[2];
    R *timg;
    long tstr;
    double h;
};

// Trace gather of one grid after a step (the batched form reads one per grid).
template <typename R> struct RecArgs {
    const R *fw;
    long fs;
    const long *rlin;
    R *seis;
    int *step;
    int nrecv, nt;
};

// Argument image of one grid in device memory: the 4 per-stage StageArgs
// blocks, the 4 per-stage TimgArgs blocks and the RecArgs block.  A batch
// lays the images of its grids end to end, so grid c's blocks sit at
// c * kCtxArgBytes from grid 0's.
constexpr int kStageArgsBytes = 1024, kTimgArgsBytes = 512, kRecArgsBytes = 64;
constexpr int kTimgArgsOffset = 4 * kStageArgsBytes;
constexpr int kRecArgsOffset = kTimgArgsOffset + 4 * kTimgArgsBytes;
constexpr int kCtxArgBytes = 8192;
static_assert(kRecArgsOffset + kRecArgsBytes <= kCtxArgBytes, "argument image too small");

// variant 0: LDS-staged xy planes (default); 1: direct loads of the in-plane
// taps; 2: split velocity/stress kernels.  `a` is the host copy of the
// argument block (launch geometry), `pa` the same block in device memory,
// read by the kernel.  tinc = 0, 1, 1, 2 in stages 0-3: the kernel forms
// Tindex = 2*(*a.step) + tinc.
template <typename R>
void launch_stage(int variant, int sx, int sz, int st, const StageArgs<R> &a, const StageArgs<R> *pa, int tinc,
                  int rows, int zchunks, hipStream_t s);
// The same stage of nctx grids in one launch of the LDS variant: grid c's
// argument block lies c * kCtxArgBytes after pa and it owns workgroups
// starts[c] .. starts[c+1]-1 (starts in device memory, starts[nctx] =
// nblocks); all grids share the step phase and the metric mode hmet.
// nctx = -n: n grids of one size and z-chunk length on one medium, the n
// copies of each tile run one after another on one XCD (starts unused).
// rows of an LDS stage-kernel tile of tw columns (64 or 32; FP64 runs 64)
int stage_tile_rows(bool fp64, int tw);
template <typename R>
void launch_stage_batch(int sx, int sz, int st, int hmet, int hybrid, int tw, const StageArgs<R> *pa,
                        const int *starts, int nctx, int nblocks, int tinc, hipStream_t s);
// traction image: of one grid with its arguments av (pa null, nctx 1), or of
// nctx grids whose blocks lie at kCtxArgBytes strides from pa (av unused);
// ni x nj covers every grid's surface, cols is their total column count
template <typename R>
void launch_timg(int sx, int sz, int ni, int nj, long cols, const TimgArgs<R> &av, const TimgArgs<R> *pa, int nctx,
                 int rows, hipStream_t s);
template <typename R>
void launch_velpd(const Geom &g, const R *met, const R *med, const R *pml, R *cov, long cstr, hipStream_t s);
template <typename R>
void launch_to_padded(const R *src, R *dst, const Geom &g, hipStream_t s);
template <typename R>
void launch_from_padded(const R *src, R *dst, const Geom &g, hipStream_t s);
template <typename R>
void launch_src_prep(const Geom &g, const R *met, const R *med, const int4 *sent, const R *fvec, const R *mvec,
                     R *scoef, int nent, R hr, hipStream_t s);
template <typename R> void launch_media_prep(const Geom &g, R *med, hipStream_t s);
// Trace and snapshot gathers after a step read the step number from the
// device counter, so a captured step graph replays at any step.
template <typename R>
void launch_record(const R *fw, long fs, const long *rlin, int nrecv, R *seis, int nt, const int *step,
                   hipStream_t s);
template <typename R>
void launch_snap(const R *fw, long fs, const long *lin, int npts, int c0, int ncomp, R *buf, const int *step,
                 int tinv, int ntime, hipStream_t s);
// step counter += 1 (one thread)
void launch_step_advance(int *step, hipStream_t s);
// The batched forms read the RecArgs blocks of nctx argument images from pa:
// trace gathers of every grid (grids without receivers return at once), and
// every grid's step counter += 1.
template <typename R> void launch_record_batch(const RecArgs<R> *pa, int nctx, int max_nrecv, hipStream_t s);
void launch_step_advance_batch(const char *pa, int nctx, hipStream_t s);

} // namespace cgfd
