// cgfd_capi.hip -- implementation of the C-ABI declared in include/cgfd3d.h.
//
// Owns every device buffer of one simulation (the role of ChildProcs,
// src/typedata.h:114-217 of the reference), converts the reference's
// [x][y][z]+ghost host arrays to the engine's padded x-fastest layout on the
// GPU, and sequences the RK stages under the reference's forward/backward
// schedule (src/simulation.cc:252-717).
#include "cgfd3d.h"
#include "cgfd_kernels.h"

#include <algorithm>
#include <atomic>
#include <cstddef>
#include <cstdint>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <utility>
#include <vector>

using namespace cgfd;

namespace {

constexpr int NSETS = 4; // FW, S1, S2, tW
constexpr int SET_FW = 0, SET_S1 = 1, SET_S2 = 2, SET_TW = 3;

long round_up(long v, long m) { return (v + m - 1) / m * m; }

} // namespace

struct cgfd_ctx {
    int prec = CGFD_FP32;
    size_t esz = 4;
    int dev = 0;
    hipStream_t stream = nullptr;
    cgfd_dims d{};
    Geom g{};
    double h = 0, rh = 0;
    long nref = 0; // NX*NY*NZ of the reference layout

    char *sets = nullptr;
    char *met = nullptr;
    char *med = nullptr;
    char *pml = nullptr;
    char *staging = nullptr;
    int nabs[6] = {0, 0, 0, 0, 0, 0};
    int npml[3] = {0, 0, 0};
    long astr[3] = {0, 0, 0};
    char *aux[3] = {nullptr, nullptr, nullptr};
    char *timg = nullptr;
    long tstr = 0;
    char *cov = nullptr;
    long cstr = 0;
    char *pgv = nullptr;
    bool pgv_on = false;
    bool have_metrics = false, have_media = false;
    bool hmet = false;     // xi/eta metrics are the constants 1/0 in every valid cell
    bool use_hmet = true;  // let the stage kernel exploit it
    int zcon = 0;          // hybrid grid: engine planes z < zcon use the Cartesian contraction
    bool simplecar = false; // the reference's SimpleCar build: xi/eta metrics taken as the identity
    // the metric and media arrays (met, med), reference-counted: contexts that
    // share one medium (cgfd_share_medium) hold the same block
    struct Medium {
        char *met = nullptr, *med = nullptr;
        std::atomic<int> refs{1};
    };
    Medium *medium = nullptr;

    // point sources: host copies until init builds the device tables
    int nfrc = 0, nmnt = 0, nstf_f = 0, nstf_m = 0;
    std::vector<int> floc, mloc;
    std::vector<char> fvec, mvec, fstf, mstf;
    // focal points (source_type 1): locations and [6][n][fnt] series
    int nfoc = 0, fnt = 0;
    double fdt = 0;
    std::vector<int> rmloc;
    std::vector<char> rmser;
    int nstf_r = 0;
    int nent = 0;
    int *col_first = nullptr;
    int4 *sent = nullptr;
    char *scoef = nullptr, *fstf_d = nullptr, *mstf_d = nullptr, *fvec_d = nullptr, *mvec_d = nullptr;
    char *rstf_d = nullptr;

    int nrecv = 0;
    std::vector<int> rloc;
    long *rlin = nullptr;
    char *seis = nullptr;

    // snapshot boxes (snap_NNN of seisplot.cpp:350-361)
    struct Snapshot {
        int spec[11]; // xs ys zs xn yn zn xi yi zi tinv cmp
        int npts = 0, ntime = 0, c0 = 0, ncomp = 0;
        long *lin = nullptr; // padded index of every point, point order k*yn*zn + m*zn + n
        char *buf = nullptr; // [ntime][ncomp][npts]
    };
    std::vector<Snapshot> snaps;

    int currT = 0, stage = 0;
    int *d_step = nullptr; // device copy of currT, read by the stage, trace and snapshot kernels
    bool inited = false;
    // cgfd_run replays captured 4-step cycles, one graph per starting phase currT % 4
    bool use_graph = false, capturing = false;
    hipGraphExec_t gexec[4] = {nullptr, nullptr, nullptr, nullptr};
    int rows = 4, zchunks = 1;
    bool zc_tuned = false; // zchunks chosen by cgfd_autotune or cgfd_set_tuning: fused batches keep it
    int tile_w = 0;        // LDS kernel tile width set by cgfd_set_tile_width (0: none)
    int tile_pref = 0;     // the width cgfd_autotune found faster than the default (0: none)
    int variant = 0; // stage kernel: 0 = LDS-staged planes, 1 = direct loads
    char *args_d = nullptr; // argument image (kCtxArgBytes: stage, traction-image and trace blocks)
    bool args_dirty = true;
    // batches led by this context (the first of cgfd_run_batch's list):
    // the argument images of the batch's grids end to end, its host copy,
    // and an event that orders the other contexts' streams around it
    char *bimg_d = nullptr;
    size_t bimg_cap = 0;
    std::vector<char> bimg_h;
    hipEvent_t bev = nullptr;

    bool prof = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_stage, ev_timg;
    double ms_stage = 0, ms_timg = 0;
    long n_stage = 0, n_timg = 0;

    std::string err;

    char *set_ptr(int s) const { return sets + (size_t)s * NF * g.fs * esz; }
    char *field_ptr(int s, int c) const { return set_ptr(s) + (size_t)c * g.fs * esz; }
    char *aux_ptr(int axis, int s) const
    {
        return aux[axis] ? aux[axis] + (size_t)s * NF * astr[axis] * esz : nullptr;
    }
};

static int fail(cgfd_ctx *c, int code, const std::string &msg)
{
    if (c) c->err = msg;
    return code;
}

#define HIPCHK(ctx, call)                                                                                      \
    do {                                                                                                       \
        hipError_t e_ = (call);                                                                                \
        if (e_ != hipSuccess)                                                                                  \
            return fail(ctx, CGFD_EHIP, std::string(#call) + ": " + hipGetErrorString(e_));                   \
    } while (0)

static int dalloc(cgfd_ctx *c, char **p, size_t bytes)
{
    *p = nullptr;
    if (bytes == 0) return 0;
    hipError_t e = hipMalloc((void **)p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(c, CGFD_ENOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes: " + hipGetErrorString(e));
    }
    // zero on the context's own stream (the guard cells of every field must
    // read as zero) and wait, so later work on any stream sees it
    e = hipMemsetAsync(*p, 0, bytes, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail(c, CGFD_EHIP, std::string("hipMemsetAsync: ") + hipGetErrorString(e));
    return 0;
}

static void dfree(char *&p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

template <typename T> static void dfree_t(T *&p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

// drop this context's reference to its metric and media arrays; the last
// holder frees them
static void release_medium(cgfd_ctx *c)
{
    if (!c->medium) { // set-up failed before the block was made
        dfree(c->met);
        dfree(c->med);
    } else if (--c->medium->refs == 0) {
        dfree(c->medium->met);
        dfree(c->medium->med);
        delete c->medium;
    }
    c->medium = nullptr;
    c->met = c->med = nullptr;
}

static int own_medium(cgfd_ctx *c)
{
    c->medium = new (std::nothrow) cgfd_ctx::Medium();
    if (!c->medium) return fail(c, CGFD_ENOMEM, "out of host memory");
    c->medium->met = c->met;
    c->medium->med = c->med;
    return 0;
}

// F/B signs of stage k of step t (simulation.cc:264-641): x and y share a
// sign that alternates with the stage parity and flips for t%4 in {2,3}; z
// alternates with the stage parity and flips for t%4 in {1,2}.
static void fb_signs(int t, int k, int *sx, int *sz)
{
    const int p = t % 4;
    *sx = (((p >= 2) ? 1 : 0) ^ (k & 1)) ? -1 : 1;
    *sz = (((p == 1 || p == 2) ? 1 : 0) ^ (k & 1)) ? -1 : 1;
}

// z chunks per column sweep: enough workgroups to fill the chip (~2048
// 256-cell tiles) and at most ~15 planes per chunk, but at least 2 planes per
// chunk.  Small grids want many short chunks (64^3: 2.88 -> 4.73
// Gcell-RKstages/s going from 8 to 2 planes); short chunks also keep the
// launch's tail short (600x600x300: 6.07 ms per stage with 2 chunks of 150
// planes, 5.59 ms with 10 of 30; with later kernels 5.338 ms at 10 chunks,
// 5.328 at 12, 5.305 at 15, 5.300 at 20 and 5.313 at 25; 200x200x150 is
// fastest at 12 chunks of 13 planes).
// Tile width of the LDS stage kernel: the one set or tuned, else 32 when a
// row of ni cells takes over 8 % more lanes in 64-column tiles than in
// 32-column ones.  Measured per stage: 80 cells (128 vs 96 lanes) 13 % and
// 200 cells (256 vs 224) 10 % faster at 32; 600 cells (640 vs 608) 0.5 %
// and 64 or 192 cells (no idle lanes) 1-16 % slower, so a 32 x 8 tile costs
// about 6 % more per lane than a 64 x 4 one.  FP64 runs 64.
static bool narrow_pays(long lanes64, long lanes32) { return lanes64 * 100 > lanes32 * 108; }
static int tile_width(const cgfd_ctx *c)
{
    if (c->prec == CGFD_FP64) return 64;
    if (c->tile_w) return c->tile_w;
    if (c->tile_pref) return c->tile_pref;
    return narrow_pays((c->g.ni + 63) / 64 * 64, (c->g.ni + 31) / 32 * 32) ? 32 : 64;
}

// LDS-kernel tiles of one z plane at tile width tw
static long plane_tiles(const cgfd_ctx *c, int tw)
{
    const int tr = stage_tile_rows(c->prec == CGFD_FP64, tw);
    return (long)((c->g.ni + tw - 1) / tw) * ((c->g.nj + tr - 1) / tr);
}

static void default_tuning(cgfd_ctx *c)
{
    c->rows = 4;
    const long tiles = plane_tiles(c, tile_width(c));
    long zc = std::max((2048 + tiles - 1) / tiles, (long)(c->g.nk + 14) / 15);
    zc = std::max(1L, std::min(zc, (long)std::max(1, c->g.nk / 2)));
    c->zchunks = (int)zc;
    c->zc_tuned = false;
    c->tile_pref = 0;
}

extern "C" {

const char *cgfd_last_error(const cgfd_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int cgfd_create(const cgfd_dims *dims, int precision, int device, cgfd_ctx **out)
{
    if (!out) return CGFD_EINVAL;
    *out = nullptr;
    if (!dims || dims->ni < 1 || dims->nj < 1 || dims->nk < 4 || dims->nt < 1 || !(dims->dh > 0) ||
        !(dims->dt > 0) || (precision != CGFD_FP32 && precision != CGFD_FP64))
        return CGFD_EINVAL;
    {
        // The stage kernels address a field with 32-bit byte offsets from its
        // base pointer, so one padded field must stay below 4 GiB (e.g. FP32
        // up to about 1000^3 cells).  Larger domains must be split.
        const long nxp = round_up(dims->ni + 3, 64), plane = nxp * (dims->nj + 3L);
        const unsigned long bytes = (unsigned long)round_up((dims->nk + 8L) * plane, 64) *
                                    (precision == CGFD_FP64 ? 8 : 4);
        if (bytes > 0xffffffffUL) {
            fprintf(stderr, "cgfd_create: a padded field of %lu bytes exceeds the 4 GiB offset range of the stage "
                            "kernels; split the domain\n", bytes);
            return CGFD_EINVAL;
        }
    }
    cgfd_ctx *c = new cgfd_ctx();
    c->prec = precision;
    c->esz = precision == CGFD_FP64 ? 8 : 4;
    c->dev = device;
    c->d = *dims;
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) {
        fprintf(stderr, "cgfd_create: hipSetDevice(%d): %s\n", device, hipGetErrorString(e));
        delete c;
        return CGFD_EHIP;
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return CGFD_EHIP;
    }
    Geom &g = c->g;
    g.ni = dims->ni;
    g.nj = dims->nj;
    g.nk = dims->nk;
    g.nxp = (int)round_up(dims->ni + 3, 64);
    g.nyp = dims->nj + 3;
    g.plane = (long)g.nxp * g.nyp;
    g.fs = round_up((long)(dims->nk + 8) * g.plane, 64);
    g.base = 4 * g.plane;
    c->nref = (long)(dims->ni + 6) * (dims->nj + 6) * (dims->nk + 6);
    if (precision == CGFD_FP32) {
        c->h = (double)(float)dims->dh;
    } else {
        c->h = dims->dh;
    }
    c->rh = 1.0 / c->h;
    c->tstr = 3L * g.nj * g.nxp;
    c->cstr = (long)g.nj * g.nxp;
    int rc = 0;
    if ((rc = dalloc(c, &c->sets, (size_t)NSETS * NF * g.fs * c->esz)) ||
        (rc = dalloc(c, &c->met, (size_t)NDRV * g.fs * c->esz)) ||
        (rc = dalloc(c, &c->med, (size_t)4 * g.fs * c->esz)) ||
        (rc = own_medium(c)) ||
        (rc = dalloc(c, &c->pml, (size_t)3 * (g.ni + g.nj + g.nk) * c->esz)) ||
        (rc = dalloc(c, &c->staging, (size_t)c->nref * c->esz)) ||
        (rc = dalloc(c, &c->timg, (size_t)9 * c->tstr * c->esz)) ||
        (rc = dalloc(c, &c->cov, (size_t)18 * c->cstr * c->esz)) ||
        (rc = dalloc(c, &c->pgv, (size_t)3 * c->cstr * c->esz)) ||
        (rc = dalloc(c, (char **)&c->d_step, sizeof(int))) || (rc = dalloc(c, &c->args_d, kCtxArgBytes))) {
        fprintf(stderr, "cgfd_create: %s\n", c->err.c_str());
        cgfd_destroy(c);
        return rc;
    }
    if (hipEventCreateWithFlags(&c->bev, hipEventDisableTiming) != hipSuccess) {
        fprintf(stderr, "cgfd_create: hipEventCreateWithFlags failed\n");
        cgfd_destroy(c);
        return CGFD_EHIP;
    }
    // unit scaling factors until cgfd_set_pml supplies tables
    if (precision == CGFD_FP32) {
        std::vector<float> t(3 * (g.ni + g.nj + g.nk), 0.0f);
        for (int i = 0; i < g.ni; i++) t[g.ni + i] = 1;
        for (int i = 0; i < g.nj; i++) t[3 * g.ni + g.nj + i] = 1;
        for (int i = 0; i < g.nk; i++) t[3 * g.ni + 3 * g.nj + g.nk + i] = 1;
        (void)hipMemcpy(c->pml, t.data(), t.size() * 4, hipMemcpyHostToDevice);
    } else {
        std::vector<double> t(3 * (g.ni + g.nj + g.nk), 0.0);
        for (int i = 0; i < g.ni; i++) t[g.ni + i] = 1;
        for (int i = 0; i < g.nj; i++) t[3 * g.ni + g.nj + i] = 1;
        for (int i = 0; i < g.nk; i++) t[3 * g.ni + 3 * g.nj + g.nk + i] = 1;
        (void)hipMemcpy(c->pml, t.data(), t.size() * 8, hipMemcpyHostToDevice);
    }
    default_tuning(c);
    *out = c;
    return 0;
}

static void free_snaps(cgfd_ctx *c)
{
    for (auto &sn : c->snaps) {
        if (sn.lin) (void)hipFree(sn.lin);
        if (sn.buf) (void)hipFree(sn.buf);
    }
    c->snaps.clear();
}

static void free_sources(cgfd_ctx *c)
{
    dfree_t(c->col_first);
    dfree_t(c->sent);
    dfree(c->scoef);
    dfree(c->fstf_d);
    dfree(c->mstf_d);
    dfree(c->fvec_d);
    dfree(c->mvec_d);
    dfree(c->rstf_d);
    c->nent = 0;
}

static void drop_graphs(cgfd_ctx *c)
{
    for (auto &g : c->gexec) {
        if (g) (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
}

static void drop_events(cgfd_ctx *c)
{
    for (auto &p : c->ev_stage) {
        (void)hipEventDestroy(p.first);
        (void)hipEventDestroy(p.second);
    }
    for (auto &p : c->ev_timg) {
        (void)hipEventDestroy(p.first);
        (void)hipEventDestroy(p.second);
    }
    c->ev_stage.clear();
    c->ev_timg.clear();
}

void cgfd_destroy(cgfd_ctx *c)
{
    if (!c) return;
    (void)hipSetDevice(c->dev);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    drop_events(c);
    drop_graphs(c);
    dfree_t(c->d_step);
    dfree(c->sets);
    release_medium(c);
    dfree(c->pml);
    dfree(c->staging);
    for (int a = 0; a < 3; a++) dfree(c->aux[a]);
    dfree(c->timg);
    dfree(c->cov);
    dfree(c->pgv);
    free_sources(c);
    dfree_t(c->rlin);
    dfree(c->seis);
    free_snaps(c);
    dfree(c->args_d);
    dfree(c->bimg_d);
    if (c->bev) (void)hipEventDestroy(c->bev);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

// upload one reference-layout array into a padded field
static int upload_field(cgfd_ctx *c, const void *src, char *dst)
{
    HIPCHK(c, hipMemcpyAsync(c->staging, src, (size_t)c->nref * c->esz, hipMemcpyHostToDevice, c->stream));
    if (c->prec == CGFD_FP32)
        launch_to_padded((const float *)c->staging, (float *)dst, c->g, c->stream);
    else
        launch_to_padded((const double *)c->staging, (double *)dst, c->g, c->stream);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

static int download_field(cgfd_ctx *c, const char *src, void *dst)
{
    HIPCHK(c, hipMemsetAsync(c->staging, 0, (size_t)c->nref * c->esz, c->stream));
    if (c->prec == CGFD_FP32)
        launch_from_padded((const float *)src, (float *)c->staging, c->g, c->stream);
    else
        launch_from_padded((const double *)src, (double *)c->staging, c->g, c->stream);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(dst, c->staging, (size_t)c->nref * c->esz, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

} // extern "C"

// True when the six xi/eta metrics of every valid cell are xix = etay = 1
// exactly and xiy = xiz = etax = etaz = 0 (+0 or -0): meshes with fixed x and
// y spacing, whose cross-product metrics carry zeros of either sign.  The
// stage kernel then loads only the zeta metrics and multiplies by +0, which
// can change only the sign of an exactly-zero intermediate, never a value.
template <typename U> static bool horizontal_metrics_const(const cgfd_ctx *c, const void *const *drv, U one, U sign)
{
    const long NY = c->g.nj + 6, NZ = c->g.nk + 6;
    const bool is_one[6] = {true, false, false, false, true, false}; // xix xiy xiz etax etay etaz
    for (int m = 0; m < 6; m++) {
        const U *a = (const U *)drv[m];
        for (long x = 3; x < c->g.ni + 3; x++)
            for (long y = 3; y < c->g.nj + 3; y++) {
                const U *col = a + (x * NY + y) * NZ + 3;
                for (long z = 0; z < c->g.nk; z++)
                    if (is_one[m] ? col[z] != one : (col[z] & ~sign) != 0) return false;
            }
    }
    return true;
}

extern "C" {

int cgfd_set_metrics(cgfd_ctx *c, const void *const *drv_in)
{
    if (!c || !drv_in) return fail(c, CGFD_EINVAL, "null argument");
    if (c->medium && c->medium->refs > 1)
        return fail(c, CGFD_EINVAL, "the metrics are shared with another context (cgfd_share_medium)");
    HIPCHK(c, hipSetDevice(c->dev));
    for (int k = 0; k < NDRV; k++)
        if (!drv_in[k]) return fail(c, CGFD_EINVAL, "null metric array");
    // SimpleCar: the xi/eta metrics are the constants 1 (xi_x, eta_y) and 0,
    // as the reference's SimpleCar build assumes them (calculate.cu:3369-3375,
    // 4339-4456, 5729-5740); with them the curvilinear and Cartesian forms
    // compute that build's values (a product with 1 and a sum with a zero are
    // exact)
    const void *drv[NDRV];
    std::vector<unsigned char> one, zero;
    for (int k = 0; k < NDRV; k++) drv[k] = drv_in[k];
    if (c->simplecar) {
        const size_t n = (size_t)(c->g.ni + 6) * (c->g.nj + 6) * (c->g.nk + 6);
        one.assign(n * c->esz, 0);
        zero.assign(n * c->esz, 0);
        for (size_t i = 0; i < n; i++) {
            if (c->prec == CGFD_FP32) ((float *)one.data())[i] = 1.0f;
            else ((double *)one.data())[i] = 1.0;
        }
        for (int k = XIX; k <= ETZ; k++) drv[k] = (k == XIX || k == ETY) ? one.data() : zero.data();
    }
    for (int k = 0; k < NDRV; k++) {
        int rc = upload_field(c, drv[k], c->met + (size_t)k * c->g.fs * c->esz);
        if (rc) return rc;
    }
    // compare bit patterns: 1.0 exactly, zeros of either sign
    c->hmet = c->prec == CGFD_FP32
                  ? horizontal_metrics_const<uint32_t>(c, drv, 0x3f800000u, 0x80000000u)
                  : horizontal_metrics_const<uint64_t>(c, drv, 0x3ff0000000000000ull, 0x8000000000000000ull);
    c->have_metrics = true;
    c->inited = false;
    c->args_dirty = true;
    return 0;
}

int cgfd_set_media(cgfd_ctx *c, const void *alpha, const void *beta, const void *rho)
{
    if (!c || !alpha || !beta || !rho) return fail(c, CGFD_EINVAL, "null argument");
    if (c->medium && c->medium->refs > 1)
        return fail(c, CGFD_EINVAL, "the media are shared with another context (cgfd_share_medium)");
    HIPCHK(c, hipSetDevice(c->dev));
    // alpha -> slot 0, beta -> slot 1, rho -> slot 3; then derive
    // lam2mu, miu, 1/rho in slots 0-2 on the GPU
    const void *src[3] = {alpha, beta, rho};
    const int slot[3] = {0, 1, 3};
    for (int k = 0; k < 3; k++) {
        int rc = upload_field(c, src[k], c->med + (size_t)slot[k] * c->g.fs * c->esz);
        if (rc) return rc;
    }
    if (c->prec == CGFD_FP32)
        launch_media_prep(c->g, (float *)c->med, c->stream);
    else
        launch_media_prep(c->g, (double *)c->med, c->stream);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->have_media = true;
    c->inited = false;
    return 0;
}

int cgfd_set_pml(cgfd_ctx *c, const int *nabs, const void *const *tabs)
{
    if (!c || !nabs || !tabs) return fail(c, CGFD_EINVAL, "null argument");
    const int nv[3] = {c->g.ni, c->g.nj, c->g.nk};
    for (int a = 0; a < 3; a++) {
        if (nabs[2 * a] < 0 || nabs[2 * a + 1] < 0)
            return fail(c, CGFD_EINVAL, "abs_number must be non-negative (absorb.cpp:151-157)");
        if (nabs[2 * a] + nabs[2 * a + 1] >= nv[a])
            return fail(c, CGFD_EINVAL, "total absorbing layers exceed the valid domain (absorb.cpp:158-163)");
    }
    for (int k = 0; k < 9; k++)
        if (!tabs[k]) return fail(c, CGFD_EINVAL, "null PML table");
    HIPCHK(c, hipSetDevice(c->dev));
    // valid-index slices of the 9 tables, ordered APD B DB per axis
    const int len[3] = {c->g.ni, c->g.nj, c->g.nk};
    std::vector<char> host((size_t)3 * (c->g.ni + c->g.nj + c->g.nk) * c->esz);
    size_t off = 0;
    for (int a = 0; a < 3; a++)
        for (int kind = 0; kind < 3; kind++) { // APD, B, DB
            const char *src = (const char *)tabs[kind * 3 + a] + 3 * c->esz;
            memcpy(host.data() + off, src, (size_t)len[a] * c->esz);
            off += (size_t)len[a] * c->esz;
        }
    HIPCHK(c, hipMemcpy(c->pml, host.data(), host.size(), hipMemcpyHostToDevice));
    memcpy(c->nabs, nabs, sizeof(c->nabs));
    for (int a = 0; a < 3; a++) dfree(c->aux[a]);
    c->npml[0] = nabs[0] + nabs[1];
    c->npml[1] = nabs[2] + nabs[3];
    c->npml[2] = nabs[4] + nabs[5];
    c->astr[0] = round_up((long)c->g.nk * c->g.nj * c->npml[0], 64);
    c->astr[1] = round_up((long)c->g.nk * c->npml[1] * c->g.nxp, 64);
    c->astr[2] = round_up((long)c->npml[2] * c->g.nj * c->g.nxp, 64);
    for (int a = 0; a < 3; a++) {
        int rc = dalloc(c, &c->aux[a], (size_t)NSETS * NF * c->astr[a] * c->esz);
        if (rc) return rc;
    }
    c->inited = false;
    return 0;
}

static int check_points(cgfd_ctx *c, int n, const int *locs)
{
    for (int i = 0; i < n; i++) {
        const int x = locs[3 * i], y = locs[3 * i + 1], z = locs[3 * i + 2];
        if (x < 3 || x >= c->g.ni + 3 || y < 3 || y >= c->g.nj + 3 || z < 3 || z >= c->g.nk + 3)
            return fail(c, CGFD_ERANGE,
                        "point " + std::to_string(i) + " outside the valid domain (source.cpp:1077-1096)");
    }
    return 0;
}

int cgfd_set_forces(cgfd_ctx *c, int n, const int *locs, const void *vec, const void *stf, int nstf)
{
    if (!c || n < 0 || (n > 0 && (!locs || !vec || !stf))) return fail(c, CGFD_EINVAL, "bad force arguments");
    if (n > 0 && nstf < 2 * c->d.nt + 1) return fail(c, CGFD_EINVAL, "force STF shorter than 2*nt+1 samples");
    int rc = check_points(c, n, locs);
    if (rc) return rc;
    c->nfrc = n;
    c->nstf_f = nstf;
    c->floc.assign(locs, locs + 3 * n);
    c->fvec.assign((const char *)vec, (const char *)vec + (size_t)3 * n * c->esz);
    c->fstf.assign((const char *)stf, (const char *)stf + (size_t)n * nstf * c->esz);
    c->inited = false;
    return 0;
}

int cgfd_set_moments(cgfd_ctx *c, int n, const int *locs, const void *vec, const void *stf, int nstf)
{
    if (!c || n < 0 || (n > 0 && (!locs || !vec || !stf))) return fail(c, CGFD_EINVAL, "bad moment arguments");
    if (n > 0 && nstf < 2 * c->d.nt + 1) return fail(c, CGFD_EINVAL, "moment STF shorter than 2*nt+1 samples");
    int rc = check_points(c, n, locs);
    if (rc) return rc;
    c->nmnt = n;
    c->nstf_m = nstf;
    c->mloc.assign(locs, locs + 3 * n);
    c->mvec.assign((const char *)vec, (const char *)vec + (size_t)6 * n * c->esz);
    c->mstf.assign((const char *)stf, (const char *)stf + (size_t)n * nstf * c->esz);
    c->inited = false;
    return 0;
}

int cgfd_set_focal(cgfd_ctx *c, int n, const int *locs, const void *series, int fnt, double fdt)
{
    if (!c || n < 0 || (n > 0 && (!locs || !series || fnt < 1 || !(fdt > 0))))
        return fail(c, CGFD_EINVAL, "bad focal arguments");
    int rc = check_points(c, n, locs);
    if (rc) return rc;
    c->nfoc = n;
    c->fnt = n > 0 ? fnt : 0;
    c->fdt = fdt;
    c->rmloc.assign(locs, locs + 3 * n);
    c->rmser.assign((const char *)series, (const char *)series + (size_t)6 * n * c->fnt * c->esz);
    c->inited = false;
    return 0;
}

int cgfd_set_receivers(cgfd_ctx *c, int n, const int *locs)
{
    if (!c || n < 0 || (n > 0 && !locs)) return fail(c, CGFD_EINVAL, "bad receiver arguments");
    int rc = check_points(c, n, locs);
    if (rc) return rc;
    HIPCHK(c, hipSetDevice(c->dev));
    c->nrecv = n;
    c->rloc.assign(locs, locs + 3 * n);
    dfree_t(c->rlin);
    dfree(c->seis);
    if (n > 0) {
        std::vector<long> lin(n);
        for (int i = 0; i < n; i++) {
            const int x = locs[3 * i] - 3, y = locs[3 * i + 1] - 3, z = locs[3 * i + 2] - 3;
            lin[i] = c->g.base + ((long)z * c->g.nyp + y) * c->g.nxp + x;
        }
        HIPCHK(c, hipMalloc((void **)&c->rlin, n * sizeof(long)));
        HIPCHK(c, hipMemcpy(c->rlin, lin.data(), n * sizeof(long), hipMemcpyHostToDevice));
        rc = dalloc(c, &c->seis, (size_t)NF * n * c->d.nt * c->esz);
        if (rc) return rc;
    }
    c->inited = false;
    return 0;
}

int cgfd_set_snapshots(cgfd_ctx *c, int nsnap, const int *spec)
{
    if (!c || nsnap < 0 || (nsnap > 0 && !spec)) return fail(c, CGFD_EINVAL, "bad snapshot arguments");
    const long NX = c->g.ni + 6, NY = c->g.nj + 6, NZ = c->g.nk + 6;
    for (int i = 0; i < nsnap; i++) {
        const int *q = spec + 11 * i;
        const std::string id = "snapshot " + std::to_string(i + 1) + ": ";
        if (q[3] < 1 || q[4] < 1 || q[5] < 1 || q[6] < 1 || q[7] < 1 || q[8] < 1)
            return fail(c, CGFD_EINVAL, id + "counts and intervals must be >= 1");
        if (q[9] < 1) return fail(c, CGFD_EINVAL, id + "tinv must be >= 1");
        if (q[10] < 1 || q[10] > 3) return fail(c, CGFD_EINVAL, id + "cmp must be 1 (V), 2 (T) or 3 (both)");
        // ghost-inclusive locations xs-1+LenFD+k*xi (seisplot.cpp:508-518)
        const long lo[3] = {q[0] + 2L, q[1] + 2L, q[2] + 2L};
        const long hi[3] = {lo[0] + (long)(q[3] - 1) * q[6], lo[1] + (long)(q[4] - 1) * q[7],
                            lo[2] + (long)(q[5] - 1) * q[8]};
        if (lo[0] < 0 || lo[1] < 0 || lo[2] < 0 || hi[0] >= NX || hi[1] >= NY || hi[2] >= NZ)
            return fail(c, CGFD_ERANGE, id + "box leaves the grid");
        if ((long)q[3] * q[4] * q[5] > (1L << 30)) return fail(c, CGFD_EINVAL, id + "too many points");
    }
    HIPCHK(c, hipSetDevice(c->dev));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    free_snaps(c);
    for (int i = 0; i < nsnap; i++) {
        cgfd_ctx::Snapshot sn;
        memcpy(sn.spec, spec + 11 * i, sizeof(sn.spec));
        const int *q = sn.spec;
        sn.npts = q[3] * q[4] * q[5];
        sn.ntime = (c->d.nt + q[9] - 1) / q[9]; // ceil(1.0*nt/tinv)
        sn.c0 = q[10] == 2 ? TXX : VX;
        sn.ncomp = q[10] == 1 ? 3 : (q[10] == 2 ? 6 : 9);
        std::vector<long> lin(sn.npts);
        for (int k = 0; k < q[3]; k++)
            for (int m = 0; m < q[4]; m++)
                for (int n = 0; n < q[5]; n++) {
                    const long x = q[0] - 1 + k * q[6], y = q[1] - 1 + m * q[7], z = q[2] - 1 + n * q[8];
                    lin[((size_t)k * q[4] + m) * q[5] + n] = c->g.base + (z * c->g.nyp + y) * c->g.nxp + x;
                }
        c->snaps.push_back(sn);
        cgfd_ctx::Snapshot &dst = c->snaps.back();
        int rc = dalloc(c, (char **)&dst.lin, lin.size() * sizeof(long));
        if (!rc) rc = dalloc(c, &dst.buf, (size_t)dst.ntime * dst.ncomp * dst.npts * c->esz);
        if (rc) {
            free_snaps(c);
            return rc;
        }
        HIPCHK(c, hipMemcpy(dst.lin, lin.data(), lin.size() * sizeof(long), hipMemcpyHostToDevice));
    }
    c->inited = false;
    return 0;
}

int cgfd_snap_shape(const cgfd_ctx *c, int isnap, int *npoints, int *ntime)
{
    if (!c || isnap < 0 || isnap >= (int)c->snaps.size()) return CGFD_EINVAL;
    if (npoints) *npoints = c->snaps[isnap].npts;
    if (ntime) *ntime = c->snaps[isnap].ntime;
    return 0;
}

int cgfd_share_medium(cgfd_ctx *c, const cgfd_ctx *src)
{
    if (!c || !src) return c ? fail(c, CGFD_EINVAL, "null argument") : CGFD_EINVAL;
    if (c == src) return fail(c, CGFD_EINVAL, "a context cannot share its own medium");
    if (src->dev != c->dev || src->prec != c->prec || src->g.ni != c->g.ni || src->g.nj != c->g.nj ||
        src->g.nk != c->g.nk)
        return fail(c, CGFD_EINVAL, "cgfd_share_medium needs a context on the same device, of the same precision "
                                    "and grid size");
    if (!src->have_metrics || !src->have_media)
        return fail(c, CGFD_EINVAL, "the source context has no metrics or media yet");
    if (c->have_metrics || c->have_media)
        return fail(c, CGFD_EINVAL, "cgfd_share_medium must come before this context's own metrics and media");
    HIPCHK(c, hipSetDevice(c->dev));
    release_medium(c);
    c->medium = src->medium;
    ++c->medium->refs;
    c->met = src->met;
    c->med = src->med;
    c->hmet = src->hmet;
    c->simplecar = src->simplecar;
    c->have_metrics = c->have_media = true;
    c->inited = false;
    c->args_dirty = true;
    return 0;
}

int cgfd_set_simplecar(cgfd_ctx *c, int on)
{
    if (!c) return CGFD_EINVAL;
    if (c->have_metrics && c->simplecar != (on != 0))
        return fail(c, CGFD_EINVAL, "cgfd_set_simplecar must precede cgfd_set_metrics (it replaces the xi/eta metrics "
                                    "as they are uploaded)");
    c->simplecar = on != 0;
    return 0;
}

int cgfd_set_hybrid(cgfd_ctx *c, int conindex)
{
    if (!c) return CGFD_EINVAL;
    if (conindex < 3 || conindex > c->g.nk)
        return fail(c, CGFD_EINVAL,
                    "conindex must lie in [3, nk]: the curvilinear contraction applies from that z index (with "
                    "the 3 ghost planes) up, and the 3 surface rows must stay curvilinear");
    c->zcon = conindex - 3;
    c->args_dirty = true;
    return 0;
}

int cgfd_set_pgv(cgfd_ctx *c, int on)
{
    if (!c) return CGFD_EINVAL;
    c->pgv_on = on != 0;
    c->args_dirty = true;
    return 0;
}

} // extern "C"

// ExtractValue (calculate.cu:2586-2613): linear interpolation of a series
// sampled every DT from Tstart, zero outside [Tstart, Tend], in the
// reference's Real arithmetic.  `avail` bounds the reads to the flat series
// array (the series of the following focal points come after this one, as in
// the reference's H_Rmnt layout).
template <typename R> static R extract_value(const R *stf, long avail, R Tstart, R Tend, R DT, R time)
{
    int PL, PR;
    R value = 0;
    if (time > Tend || time < Tstart)
        value = 0;
    else {
        if (std::fabs(time - DT * std::floor(time / DT)) < 1e-5) {
            PL = (int)(time / DT);
            PR = PL;
        } else {
            PL = (int)(time / DT);
            PR = PL + 1;
        }
        const R vl = PL < avail ? stf[PL] : (R)0, vr = PR < avail ? stf[PR] : (R)0;
        value = vl + (vr - vl) * ((time - DT * PL) / DT);
    }
    return value;
}

// The moment-rate tensor of every focal point at every stage time, indexed
// like the source time functions by Tindex = 2*currT + tinc: the stage time
// (currT + tinc*0.5)*stept (calculate.cu:1422) is a function of Tindex alone.
// Layout [point][6][2*nt+1]; InterpFocus (calculate.cu:2615-2676) computes
// the same values on the host once per stage.
template <typename R> static std::vector<R> focal_table(const cgfd_ctx *c, int nstf)
{
    const long all = (long)c->nfoc * c->fnt;
    const R *ser = (const R *)c->rmser.data();
    const R stept = (R)c->d.dt, DT = (R)c->fdt;
    const R Tstart = 0, Tend = DT * (c->fnt - 1);
    std::vector<R> tab((size_t)c->nfoc * 6 * nstf);
    for (int t = 0; t < nstf; t++) {
        const R time = (R)((t * 0.5) * stept);
        for (int i = 0; i < c->nfoc; i++) {
            const long src = (long)i * c->fnt;
            for (int m = 0; m < 6; m++)
                tab[((size_t)i * 6 + m) * nstf + t] = extract_value<R>(ser + m * all + src, all - src, Tstart, Tend, DT, time);
        }
    }
    return tab;
}

extern "C" {

// Source entries sorted by (column, z, kind, source index), so that a lane
// meets the sources of its column in z order and applies coincident ones in
// the reference's order (LoadForce, LoadMoment, LoadRmom, calculate.cu:
// 1534-1549, each in source-index order).
static int build_sources(cgfd_ctx *c)
{
    free_sources(c);
    struct Ent {
        int col, z, kind, sid;
    };
    std::vector<Ent> ents;
    for (int i = 0; i < c->nfrc; i++)
        ents.push_back({(c->floc[3 * i + 1] - 3) * c->g.ni + (c->floc[3 * i] - 3), c->floc[3 * i + 2] - 3, 0, i});
    for (int i = 0; i < c->nmnt; i++)
        ents.push_back({(c->mloc[3 * i + 1] - 3) * c->g.ni + (c->mloc[3 * i] - 3), c->mloc[3 * i + 2] - 3, 1, i});
    for (int i = 0; i < c->nfoc; i++)
        ents.push_back({(c->rmloc[3 * i + 1] - 3) * c->g.ni + (c->rmloc[3 * i] - 3), c->rmloc[3 * i + 2] - 3, 2, i});
    if (ents.empty()) return 0;
    std::stable_sort(ents.begin(), ents.end(), [](const Ent &a, const Ent &b) {
        if (a.col != b.col) return a.col < b.col;
        if (a.z != b.z) return a.z < b.z;
        if (a.kind != b.kind) return a.kind < b.kind;
        return a.sid < b.sid;
    });
    const int n = (int)ents.size();
    std::vector<int4> hs(n + 1);
    std::vector<int> first((size_t)c->g.ni * c->g.nj, -1);
    for (int e = 0; e < n; e++) {
        hs[e] = make_int4(ents[e].z, ents[e].kind, ents[e].sid, ents[e].col);
        if (first[ents[e].col] < 0) first[ents[e].col] = e;
    }
    hs[n] = make_int4(-1, -1, -1, -1);
    HIPCHK(c, hipMalloc((void **)&c->sent, (n + 1) * sizeof(int4)));
    HIPCHK(c, hipMemcpy(c->sent, hs.data(), (n + 1) * sizeof(int4), hipMemcpyHostToDevice));
    HIPCHK(c, hipMalloc((void **)&c->col_first, first.size() * sizeof(int)));
    HIPCHK(c, hipMemcpy(c->col_first, first.data(), first.size() * sizeof(int), hipMemcpyHostToDevice));
    int rc;
    if ((rc = dalloc(c, &c->scoef, (size_t)n * 8 * c->esz))) return rc;
    if (c->nfrc) {
        if ((rc = dalloc(c, &c->fvec_d, c->fvec.size())) || (rc = dalloc(c, &c->fstf_d, c->fstf.size()))) return rc;
        HIPCHK(c, hipMemcpy(c->fvec_d, c->fvec.data(), c->fvec.size(), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(c->fstf_d, c->fstf.data(), c->fstf.size(), hipMemcpyHostToDevice));
    }
    if (c->nmnt) {
        if ((rc = dalloc(c, &c->mvec_d, c->mvec.size())) || (rc = dalloc(c, &c->mstf_d, c->mstf.size()))) return rc;
        HIPCHK(c, hipMemcpy(c->mvec_d, c->mvec.data(), c->mvec.size(), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(c->mstf_d, c->mstf.data(), c->mstf.size(), hipMemcpyHostToDevice));
    }
    if (c->nfoc) {
        c->nstf_r = 2 * c->d.nt + 1;
        std::vector<char> tab;
        if (c->prec == CGFD_FP32) {
            std::vector<float> t = focal_table<float>(c, c->nstf_r);
            tab.assign((const char *)t.data(), (const char *)(t.data() + t.size()));
        } else {
            std::vector<double> t = focal_table<double>(c, c->nstf_r);
            tab.assign((const char *)t.data(), (const char *)(t.data() + t.size()));
        }
        if ((rc = dalloc(c, &c->rstf_d, tab.size()))) return rc;
        HIPCHK(c, hipMemcpy(c->rstf_d, tab.data(), tab.size(), hipMemcpyHostToDevice));
    }
    c->nent = n;
    if (c->prec == CGFD_FP32)
        launch_src_prep(c->g, (const float *)c->met, (const float *)c->med, c->sent, (const float *)c->fvec_d,
                        (const float *)c->mvec_d, (float *)c->scoef, n, (float)c->h, c->stream);
    else
        launch_src_prep(c->g, (const double *)c->met, (const double *)c->med, c->sent, (const double *)c->fvec_d,
                        (const double *)c->mvec_d, (double *)c->scoef, n, (double)c->h, c->stream);
    HIPCHK(c, hipGetLastError());
    return 0;
}

int cgfd_init(cgfd_ctx *c)
{
    if (!c) return CGFD_EINVAL;
    if (!c->have_metrics || !c->have_media) return fail(c, CGFD_ESTATE, "metrics and media must be set before init");
    HIPCHK(c, hipSetDevice(c->dev));
    HIPCHK(c, hipMemsetAsync(c->sets, 0, (size_t)NSETS * NF * c->g.fs * c->esz, c->stream));
    for (int a = 0; a < 3; a++)
        if (c->aux[a]) HIPCHK(c, hipMemsetAsync(c->aux[a], 0, (size_t)NSETS * NF * c->astr[a] * c->esz, c->stream));
    HIPCHK(c, hipMemsetAsync(c->pgv, 0, (size_t)3 * c->cstr * c->esz, c->stream));
    HIPCHK(c, hipMemsetAsync(c->timg, 0, (size_t)9 * c->tstr * c->esz, c->stream));
    HIPCHK(c, hipMemsetAsync(c->d_step, 0, sizeof(int), c->stream));
    drop_graphs(c);
    if (c->seis) HIPCHK(c, hipMemsetAsync(c->seis, 0, (size_t)NF * c->nrecv * c->d.nt * c->esz, c->stream));
    for (auto &sn : c->snaps)
        HIPCHK(c, hipMemsetAsync(sn.buf, 0, (size_t)sn.ntime * sn.ncomp * sn.npts * c->esz, c->stream));
    if (c->prec == CGFD_FP32)
        launch_velpd(c->g, (const float *)c->met, (const float *)c->med, (const float *)c->pml, (float *)c->cov,
                     c->cstr, c->stream);
    else
        launch_velpd(c->g, (const double *)c->met, (const double *)c->med, (const double *)c->pml,
                     (double *)c->cov, c->cstr, c->stream);
    HIPCHK(c, hipGetLastError());
    int rc = build_sources(c);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->currT = 0;
    c->stage = 0;
    c->inited = true;
    c->args_dirty = true;
    return 0;
}

// input / output buffer sets of stage k: FW -> S1 -> S2 -> S1 -> FW
static int stage_in(int k) { return k == 0 ? SET_FW : (k == 2 ? SET_S2 : SET_S1); }
static int stage_out(int k) { return k == 0 ? SET_S1 : (k == 1 ? SET_S2 : (k == 2 ? SET_S1 : SET_FW)); }

static void fold_events(cgfd_ctx *c)
{
    auto fold = [](std::vector<std::pair<hipEvent_t, hipEvent_t>> &v, double &ms, long &n) {
        for (auto &p : v) {
            float t = 0;
            if (hipEventSynchronize(p.second) == hipSuccess && hipEventElapsedTime(&t, p.first, p.second) == hipSuccess) {
                ms += t;
                n++;
            }
            (void)hipEventDestroy(p.first);
            (void)hipEventDestroy(p.second);
        }
        v.clear();
    };
    fold(c->ev_stage, c->ms_stage, c->n_stage);
    fold(c->ev_timg, c->ms_timg, c->n_timg);
}

} // extern "C"

template <typename R> static StageArgs<R> make_stage_args(const cgfd_ctx *c, int k)
{
    static const R RK4A[3] = {(R)0.5, (R)0.5, (R)1.0};                                      // calculate.cu:12
    static const R RK4B[4] = {(R)(1.0 / 6.0), (R)(1.0 / 3.0), (R)(1.0 / 3.0), (R)(1.0 / 6.0)}; // calculate.cu:13
    const R dt = (R)c->d.dt;
    const int sin_ = stage_in(k), sout = stage_out(k);
    StageArgs<R> sa{};
    sa.g = c->g;
    sa.win = (const R *)c->set_ptr(sin_);
    sa.fw = (const R *)c->set_ptr(SET_FW);
    sa.wout = (R *)c->set_ptr(sout);
    sa.tw = (R *)c->set_ptr(SET_TW);
    sa.met = (const R *)c->met;
    sa.med = (const R *)c->med;
    sa.pml = (const R *)c->pml;
    memcpy(sa.nabs, c->nabs, sizeof(sa.nabs));
    for (int a = 0; a < 3; a++) {
        sa.npml[a] = c->npml[a];
        sa.astr[a] = c->astr[a];
        sa.ain[a] = (const R *)c->aux_ptr(a, sin_);
        sa.af[a] = (const R *)c->aux_ptr(a, SET_FW);
        sa.aout[a] = (R *)c->aux_ptr(a, sout);
        sa.at[a] = (R *)c->aux_ptr(a, SET_TW);
    }
    sa.timg = (const R *)c->timg;
    sa.tstr = c->tstr;
    sa.cov = (const R *)c->cov;
    sa.cstr = c->cstr;
    sa.col_first = c->nent ? c->col_first : nullptr;
    sa.sent = c->sent;
    sa.scoef = (const R *)c->scoef;
    sa.fstf = (const R *)c->fstf_d;
    sa.mstf = (const R *)c->mstf_d;
    sa.rstf = (const R *)c->rstf_d;
    sa.nstf_f = c->nstf_f;
    sa.nstf_m = c->nstf_m;
    sa.nstf_r = c->nstf_r;
    sa.pgv = (k == 3 && c->pgv_on) ? (R *)c->pgv : nullptr;
    sa.h = c->h;
    sa.rh = c->rh;
    sa.hr = (R)c->h;
    sa.alpha = k < 3 ? RK4A[k] * dt : (R)0;
    sa.beta = RK4B[k] * dt;
    sa.zlen = (c->g.nk + c->zchunks - 1) / c->zchunks;
    sa.hmet = (c->hmet && c->use_hmet) ? 1 : 0;
    sa.zcon = c->zcon;
    sa.tile = tile_width(c);
    sa.step = c->d_step;
    return sa;
}

template <typename R> static TimgArgs<R> make_timg_args(const cgfd_ctx *c, int k)
{
    const int sin_ = stage_in(k);
    TimgArgs<R> ta{};
    ta.g = c->g;
    ta.win = (const R *)c->set_ptr(sin_);
    ta.met = (const R *)c->met;
    ta.med = (const R *)c->med;
    ta.pml = (const R *)c->pml;
    memcpy(ta.nabs, c->nabs, sizeof(ta.nabs));
    for (int a = 0; a < 2; a++) {
        ta.npml[a] = c->npml[a];
        ta.astr[a] = c->astr[a];
        ta.ain[a] = (const R *)c->aux_ptr(a, sin_);
    }
    ta.timg = (R *)c->timg;
    ta.tstr = c->tstr;
    ta.h = c->h;
    return ta;
}

template <typename R> static RecArgs<R> make_rec_args(const cgfd_ctx *c)
{
    RecArgs<R> ra{};
    ra.fw = (const R *)c->set_ptr(SET_FW);
    ra.fs = c->g.fs;
    ra.rlin = c->rlin;
    ra.seis = (R *)c->seis;
    ra.step = c->d_step;
    ra.nrecv = c->nrecv;
    ra.nt = c->d.nt;
    return ra;
}

// the argument image of one grid (layout in cgfd_kernels.h); zlen overrides
// the stage blocks' z-chunk length when positive
template <typename R> static void make_image(const cgfd_ctx *c, char *dst, int zlen)
{
    static_assert(sizeof(StageArgs<R>) <= kStageArgsBytes && sizeof(TimgArgs<R>) <= kTimgArgsBytes &&
                      sizeof(RecArgs<R>) <= kRecArgsBytes,
                  "argument block too large");
    static_assert(offsetof(RecArgs<float>, step) == offsetof(RecArgs<double>, step), "step counter offset");
    memset(dst, 0, kCtxArgBytes);
    for (int k = 0; k < 4; k++) {
        StageArgs<R> a = make_stage_args<R>(c, k);
        if (zlen > 0) a.zlen = zlen;
        memcpy(dst + k * kStageArgsBytes, &a, sizeof(a));
        TimgArgs<R> t = make_timg_args<R>(c, k);
        memcpy(dst + kTimgArgsOffset + k * kTimgArgsBytes, &t, sizeof(t));
    }
    RecArgs<R> r = make_rec_args<R>(c);
    memcpy(dst + kRecArgsOffset, &r, sizeof(r));
}

// (re)write the context's argument image
static int upload_args(cgfd_ctx *c)
{
    std::vector<char> host(kCtxArgBytes, 0);
    if (c->prec == CGFD_FP32) make_image<float>(c, host.data(), 0);
    else make_image<double>(c, host.data(), 0);
    if (!c->args_d) {
        int rc = dalloc(c, &c->args_d, host.size());
        if (rc) return rc;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(c->args_d, host.data(), host.size(), hipMemcpyHostToDevice));
    c->args_dirty = false;
    drop_graphs(c); // captured launches refer to the old blocks' contents
    return 0;
}

template <typename R> static int run_stage(cgfd_ctx *c)
{
    const int k = c->stage;
    int sx, sz;
    fb_signs(c->currT, k, &sx, &sz);
    const int tinc = k == 0 ? 0 : (k < 3 ? 1 : 2); // Tindex = 2*currT + tinc, calculate.cu:1404-1421
    if (c->args_dirty) {
        int rc = upload_args(c);
        if (rc) return rc;
    }
    const StageArgs<R> sa = make_stage_args<R>(c, k);
    const StageArgs<R> *pa = (const StageArgs<R> *)(c->args_d + k * kStageArgsBytes);
    const TimgArgs<R> ta = make_timg_args<R>(c, k);

    // profiling: events bracket each launch on the context's stream
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    if (c->prof) {
        for (auto &e : ev) HIPCHK(c, hipEventCreate(&e));
        HIPCHK(c, hipEventRecord(ev[0], c->stream));
    }
    launch_timg<R>(sx, sz, c->g.ni, c->g.nj, (long)c->g.ni * c->g.nj, ta, nullptr, 1, c->rows, c->stream);
    if (c->prof) {
        HIPCHK(c, hipEventRecord(ev[1], c->stream));
        HIPCHK(c, hipEventRecord(ev[2], c->stream));
    }
    launch_stage<R>(c->variant, sx, sz, k, sa, pa, tinc, c->rows, c->zchunks, c->stream);
    if (c->prof) {
        HIPCHK(c, hipEventRecord(ev[3], c->stream));
        c->ev_timg.push_back({ev[0], ev[1]});
        c->ev_stage.push_back({ev[2], ev[3]});
    }
    HIPCHK(c, hipGetLastError());
    if (k == 3) {
        if (c->currT < c->d.nt && c->nrecv)
            launch_record<R>((const R *)c->set_ptr(SET_FW), c->g.fs, c->rlin, c->nrecv, (R *)c->seis, c->d.nt,
                             c->d_step, c->stream);
        // SWpick after the step (simulation.cc:312-315): every tinv steps.  The
        // kernel checks the step itself; a captured cycle launches it every step.
        for (auto &sn : c->snaps) {
            const int tinv = sn.spec[9];
            if (!c->capturing && (c->currT % tinv != 0 || c->currT / tinv >= sn.ntime)) continue;
            launch_snap<R>((const R *)c->set_ptr(SET_FW), c->g.fs, sn.lin, sn.npts, sn.c0, sn.ncomp, (R *)sn.buf,
                           c->d_step, tinv, sn.ntime, c->stream);
        }
        launch_step_advance(c->d_step, c->stream);
        HIPCHK(c, hipGetLastError());
        c->currT++;
        c->stage = 0;
    } else {
        c->stage = k + 1;
    }
    if (c->prof && c->ev_stage.size() >= 1024) fold_events(c);
    return 0;
}

extern "C" {

int cgfd_run_stages(cgfd_ctx *c, int nstages)
{
    if (!c || nstages < 0) return CGFD_EINVAL;
    if (!c->inited) return fail(c, CGFD_ESTATE, "cgfd_init must run after the last set_* call");
    HIPCHK(c, hipSetDevice(c->dev));
    if ((long)c->currT * 4 + c->stage + nstages > 4L * c->d.nt)
        return fail(c, CGFD_ESTATE, "cannot step past nt (the source tables and traces are sized for nt steps)");
    for (int s = 0; s < nstages; s++) {
        int rc = c->prec == CGFD_FP32 ? run_stage<float>(c) : run_stage<double>(c);
        if (rc) return rc;
    }
    return 0;
}

// Capture the 16 stages of the 4 steps from the current one (phase currT % 4
// fixes their forward/backward pattern) into a graph.  Every step-dependent
// value reaches the kernels through the device step counter, so the graph
// replays at any step of the same phase.
static int capture_cycle(cgfd_ctx *c, int phase)
{
    if (c->args_dirty) {
        int rc = upload_args(c); // outside the capture: a copy must not become a graph node
        if (rc) return rc;
    }
    const int t0 = c->currT;
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
    c->capturing = true;
    int rc = 0;
    for (int s = 0; s < 16 && !rc; s++) rc = c->prec == CGFD_FP32 ? run_stage<float>(c) : run_stage<double>(c);
    c->capturing = false;
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(c->stream, &graph);
    c->currT = t0;
    c->stage = 0;
    if (rc || e != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc ? rc : fail(c, CGFD_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    }
    e = hipGraphInstantiate(&c->gexec[phase], graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
        c->gexec[phase] = nullptr;
        return fail(c, CGFD_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    return 0;
}

// Replay (capturing first if needed) the 4-step cycle that starts at the
// current step, on the context's stream.
static int launch_cycle(cgfd_ctx *c)
{
    const int phase = c->currT % 4;
    if (c->args_dirty) {
        int rc = upload_args(c);
        if (rc) return rc;
    }
    if (!c->gexec[phase]) {
        int rc = capture_cycle(c, phase);
        if (rc) return rc;
    }
    HIPCHK(c, hipGraphLaunch(c->gexec[phase], c->stream));
    c->currT += 4;
    return 0;
}

int cgfd_run(cgfd_ctx *c, int nsteps)
{
    if (!c || nsteps < 0) return CGFD_EINVAL;
    if (c->stage != 0) return fail(c, CGFD_ESTATE, "cgfd_run called in the middle of a step");
    if (c->use_graph && !c->prof && nsteps >= 4) {
        if (!c->inited) return fail(c, CGFD_ESTATE, "cgfd_init must run after the last set_* call");
        if ((long)c->currT + nsteps > c->d.nt)
            return fail(c, CGFD_ESTATE, "cannot step past nt (the source tables and traces are sized for nt steps)");
        HIPCHK(c, hipSetDevice(c->dev));
        for (; nsteps >= 4; nsteps -= 4) {
            int rc = launch_cycle(c);
            if (rc) return rc;
        }
    }
    return cgfd_run_stages(c, 4 * nsteps);
}

// Graph replay of whole 4-step cycles in cgfd_run (no reference
// counterpart): for small grids whose steps are bound by launch gaps.
int cgfd_set_graph(cgfd_ctx *c, int on)
{
    if (!c) return CGFD_EINVAL;
    c->use_graph = on != 0;
    if (!c->use_graph) drop_graphs(c);
    return 0;
}

// Many independent simulations advanced together: stage k of every context
// is enqueued on that context's own stream before stage k+1 of any, so the
// GPU overlaps small grids that would leave most CUs idle on their own, and
// one context's launch gaps are filled by the others' kernels.
} // extern "C"

// A batch runs as one launch per kernel when its grids share a device, the
// precision, the LDS stage variant with the same metric mode, and the step
// phase currT % 4 (which fixes the forward/backward pattern of every
// stage), and no context is timing its kernels.  Grid sizes may differ.

static bool batch_fusable(cgfd_ctx *const *ctxs, int n)
{
    if (n < 2 || n > 4096) return false; // 9 traction-image parts per grid in a 65535 grid.z
    const cgfd_ctx *c0 = ctxs[0];
    auto hm = [](const cgfd_ctx *c) { return c->hmet && c->use_hmet; };
    long most = 0; // workgroups at the most z chunks batch_layout can pick
    for (int i = 0; i < n; i++) {
        const cgfd_ctx *c = ctxs[i];
        if (c->prof || c->variant != 0) return false;
        most += plane_tiles(c, tile_width(c)) * std::max({1, c->g.nk / 2, c->zc_tuned ? c->zchunks : 0});
        if (i == 0) continue;
        if (c->dev != c0->dev || c->prec != c0->prec || c->currT % 4 != c0->currT % 4 || hm(c) != hm(c0))
            return false;
        for (int j = 0; j < i; j++)
            if (ctxs[j] == c) return false; // a grid listed twice would be stepped twice per launch
    }
    return most < 0x7fffffffL;
}

// Workgroups of a batched stage launch (the grids of one tile width tw):
// every grid sweeps its columns in about the same number of z chunks, chosen
// for about CGFD_BATCH_WORKGROUPS workgroups over the launch, with at most
// ~25 and at least 2 planes per chunk per grid; a grid whose z-chunk count
// was tuned (cgfd_autotune, cgfd_set_tuning) keeps its own.  zlen[c] is grid
// c's chunk length, starts[c] its first workgroup.
#ifndef CGFD_BATCH_WORKGROUPS
#define CGFD_BATCH_WORKGROUPS 2048
#endif
#ifndef CGFD_BATCH_PLANES
#define CGFD_BATCH_PLANES 25 // at most about this many planes per chunk
#endif
static void batch_layout(cgfd_ctx *const *ctxs, int n, int tw, int *zlen, int *starts)
{
    long tiles = 0;
    for (int i = 0; i < n; i++) tiles += plane_tiles(ctxs[i], tw);
    const long zc = std::max(1L, (CGFD_BATCH_WORKGROUPS + tiles - 1) / tiles);
    starts[0] = 0;
    for (int i = 0; i < n; i++) {
        const int nk = ctxs[i]->g.nk;
        const int zci = ctxs[i]->zc_tuned ? std::max(1, std::min(ctxs[i]->zchunks, nk))
                                          : (int)std::min(std::max(zc, (long)(nk + CGFD_BATCH_PLANES - 1) /
                                                                              CGFD_BATCH_PLANES),
                                                          (long)std::max(1, nk / 2));
        zlen[i] = (nk + zci - 1) / zci;
        starts[i + 1] = starts[i] + (int)(plane_tiles(ctxs[i], tw) * zci);
    }
}

// The device side of the allocation-free batch entry point: one fixed
// array per device for the argument images of up to kStaticBatchMax grids
// and their workgroup starts.  Host code reaches it only through the
// address hipGetSymbolAddress returns for the current device.
constexpr int kStaticBatchMax = 64;
constexpr size_t kStaticBatchBytes =
    ((size_t)kStaticBatchMax * kCtxArgBytes + (kStaticBatchMax + 2) * sizeof(int) + 255) / 256 * 256;
static __device__ __align__(256) char scratch[kStaticBatchBytes];

namespace {
constexpr int kMaxDevices = 64;
struct ScratchDev {
    bool ready = false;
    char *ptr = nullptr;       // device address of `scratch` on this device
    hipEvent_t done = nullptr; // recorded after the last launch that read it
    bool pending = false;
};
ScratchDev g_scratch[kMaxDevices];
std::mutex g_scratch_mu;
} // namespace

// per-device set-up on the first call: the array's device address and the
// event that orders rewrites of it after the previous call's launches
static int scratch_setup(cgfd_ctx *c, ScratchDev **out)
{
    if (c->dev < 0 || c->dev >= kMaxDevices) return fail(c, CGFD_EINVAL, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    ScratchDev &sd = g_scratch[c->dev];
    if (!sd.ready) {
        void *ptr = nullptr;
        HIPCHK(c, hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        HIPCHK(c, hipEventCreateWithFlags(&sd.done, hipEventDisableTiming));
        sd.ptr = (char *)ptr;
        sd.ready = true;
    }
    *out = &sd;
    return 0;
}

template <typename R> static int run_batch_fused(cgfd_ctx *const *ctxs_in, int n, int nsteps, ScratchDev *fixed)
{
    cgfd_ctx *c0 = ctxs_in[0];
    HIPCHK(c0, hipSetDevice(c0->dev));
    // Every grid runs at its own tile width: the grids of width 64 first,
    // then those of width 32 (group 1), one stage launch per group.  The
    // workgroup starts of each group begin at 0: group g's table lies at
    // starts + gfirst[g] + g.
    cgfd_ctx *ord_fixed[kStaticBatchMax];
    int zlen[kStaticBatchMax], starts_fixed[kStaticBatchMax + 2];
    std::vector<cgfd_ctx *> ov;
    std::vector<int> zv, sv;
    cgfd_ctx **ctxs = ord_fixed;
    int *zl = zlen, *starts = starts_fixed;
    if (!fixed) {
        for (int i = 0; i < n; i++)
            if (ctxs_in[i]->args_dirty) {
                int rc = upload_args(ctxs_in[i]);
                if (rc) return rc;
            }
        ov.resize(n);
        zv.resize(n);
        sv.resize(n + 2);
        ctxs = ov.data();
        zl = zv.data();
        starts = sv.data();
    }
    int gcount[2] = {0, 0};
    for (int i = 0; i < n; i++)
        if (tile_width(ctxs_in[i]) == 64) ctxs[gcount[0]++] = ctxs_in[i];
    for (int i = 0; i < n; i++)
        if (tile_width(ctxs_in[i]) != 64) ctxs[gcount[0] + gcount[1]++] = ctxs_in[i];
    const int gfirst[2] = {0, gcount[0]}, gwidth[2] = {64, 32};
    bool interleave[2] = {false, false};
    for (int g = 0; g < 2; g++) {
        if (!gcount[g]) continue;
        cgfd_ctx *const *gc = ctxs + gfirst[g];
        int *gs = starts + gfirst[g] + g, *gz = zl + gfirst[g];
        const int m = gcount[g];
        batch_layout(gc, m, gwidth[g], gz, gs);
        // grids that share one medium and have the same size and chunks run
        // interleaved (the copies of one tile together on one XCD) and read
        // the shared metric and media lines from its L2
        interleave[g] = m > 1;
        for (int i = 1; i < m && interleave[g]; i++)
            interleave[g] = gc[i]->met == gc[0]->met && gc[i]->g.ni == gc[0]->g.ni && gc[i]->g.nj == gc[0]->g.nj &&
                            gc[i]->g.nk == gc[0]->g.nk && gz[i] == gz[0] && gs[i + 1] - gs[i] == gs[1] - gs[0];
    }
    const size_t sofs = (size_t)n * kCtxArgBytes, sbytes = (size_t)(n + 2) * sizeof(int);
    int max_nrecv = 0, ni = 0, nj = 0;
    long cols = 0;
    for (int i = 0; i < n; i++) {
        const cgfd_ctx *c = ctxs[i];
        max_nrecv = std::max(max_nrecv, c->nrecv);
        ni = std::max(ni, c->g.ni);
        nj = std::max(nj, c->g.nj);
        cols += (long)c->g.ni * c->g.nj;
    }
    const char *img_d = nullptr;
    if (fixed) {
        // the previous call's launches may still read the array; then each
        // grid's image goes over from a stack buffer (nothing is allocated)
        if (fixed->pending) HIPCHK(c0, hipEventSynchronize(fixed->done));
        fixed->pending = false;
        alignas(16) char one[kCtxArgBytes];
        for (int i = 0; i < n; i++) {
            make_image<R>(ctxs[i], one, zl[i]);
            HIPCHK(c0, hipMemcpy(fixed->ptr + (size_t)i * kCtxArgBytes, one, kCtxArgBytes, hipMemcpyHostToDevice));
        }
        HIPCHK(c0, hipMemcpy(fixed->ptr + sofs, starts, sbytes, hipMemcpyHostToDevice));
        img_d = fixed->ptr;
    } else {
        // device image: the grids' argument images, then the workgroup starts
        std::vector<char> img(sofs + sbytes);
        for (int i = 0; i < n; i++) make_image<R>(ctxs[i], img.data() + (size_t)i * kCtxArgBytes, zl[i]);
        memcpy(img.data() + sofs, starts, sbytes);
        if (img != c0->bimg_h) {
            // launches of an earlier batch on c0's stream may still read the image
            HIPCHK(c0, hipStreamSynchronize(c0->stream));
            if (c0->bimg_cap < img.size()) {
                dfree(c0->bimg_d);
                c0->bimg_cap = 0;
                int rc = dalloc(c0, &c0->bimg_d, img.size());
                if (rc) return rc;
                c0->bimg_cap = img.size();
            }
            HIPCHK(c0, hipMemcpy(c0->bimg_d, img.data(), img.size(), hipMemcpyHostToDevice));
            c0->bimg_h.swap(img);
        }
        img_d = c0->bimg_d;
    }
    // c0's stream runs the batch: it waits for work queued on the others
    for (int i = 1; i < n; i++) {
        cgfd_ctx *c = ctxs_in[i];
        HIPCHK(c, hipEventRecord(c->bev, c->stream));
        HIPCHK(c0, hipStreamWaitEvent(c0->stream, c->bev, 0));
    }
    const int *starts_d = (const int *)(img_d + sofs);
    const int hmet = c0->hmet && c0->use_hmet ? 1 : 0;
    int hybrid = 0; // one grid with Cartesian planes selects the kernel that has the branch
    for (int i = 0; i < n; i++) hybrid |= ctxs[i]->zcon > 0 ? 1 : 0;
    for (int t = 0; t < nsteps; t++) {
        for (int k = 0; k < 4; k++) {
            int sx, sz;
            fb_signs(c0->currT, k, &sx, &sz);
            const int tinc = k == 0 ? 0 : (k < 3 ? 1 : 2);
            launch_timg<R>(sx, sz, ni, nj, cols, TimgArgs<R>{},
                           (const TimgArgs<R> *)(img_d + kTimgArgsOffset + k * kTimgArgsBytes), n, 4, c0->stream);
            for (int g = 0; g < 2; g++) {
                if (!gcount[g]) continue;
                const int fo = gfirst[g] + g; // this group's starts table
                launch_stage_batch<R>(sx, sz, k, hmet, hybrid, gwidth[g],
                                      (const StageArgs<R> *)(img_d + (size_t)gfirst[g] * kCtxArgBytes +
                                                             k * kStageArgsBytes),
                                      starts_d + fo, interleave[g] ? -gcount[g] : gcount[g], starts[fo + gcount[g]],
                                      tinc, c0->stream);
            }
        }
        launch_record_batch<R>((const RecArgs<R> *)(img_d + kRecArgsOffset), n, max_nrecv, c0->stream);
        for (int i = 0; i < n; i++) {
            cgfd_ctx *c = ctxs[i];
            for (auto &sn : c->snaps) {
                const int tinv = sn.spec[9];
                if (c->currT % tinv != 0 || c->currT / tinv >= sn.ntime) continue;
                launch_snap<R>((const R *)c->set_ptr(SET_FW), c->g.fs, sn.lin, sn.npts, sn.c0, sn.ncomp,
                               (R *)sn.buf, c->d_step, tinv, sn.ntime, c0->stream);
            }
        }
        launch_step_advance_batch(img_d + kRecArgsOffset, n, c0->stream);
        HIPCHK(c0, hipGetLastError());
        for (int i = 0; i < n; i++) ctxs[i]->currT++;
    }
    // the other streams wait for the batch before their own later work
    HIPCHK(c0, hipEventRecord(c0->bev, c0->stream));
    for (int i = 1; i < n; i++) HIPCHK(ctxs_in[i], hipStreamWaitEvent(ctxs_in[i]->stream, c0->bev, 0));
    if (fixed) {
        HIPCHK(c0, hipEventRecord(fixed->done, c0->stream));
        fixed->pending = true;
    }
    return 0;
}

extern "C" {

int cgfd_batch_fusable(cgfd_ctx *const *ctxs, int n)
{
    if (!ctxs || n < 0) return CGFD_EINVAL;
    for (int i = 0; i < n; i++)
        if (!ctxs[i]) return CGFD_EINVAL;
    return batch_fusable(ctxs, n) ? 1 : 0;
}

int cgfd_run_batch(cgfd_ctx *const *ctxs, int n, int nsteps)
{
    if (!ctxs || n < 0 || nsteps < 0) return CGFD_EINVAL;
    for (int i = 0; i < n; i++) {
        cgfd_ctx *c = ctxs[i];
        if (!c) return CGFD_EINVAL;
        if (!c->inited) return fail(c, CGFD_ESTATE, "cgfd_init must run after the last set_* call");
        if (c->stage != 0) return fail(c, CGFD_ESTATE, "cgfd_run_batch called in the middle of a step");
        if ((long)c->currT + nsteps > c->d.nt)
            return fail(c, CGFD_ESTATE, "cannot step past nt (the source tables and traces are sized for nt steps)");
    }
    if (nsteps > 0 && batch_fusable(ctxs, n))
        return ctxs[0]->prec == CGFD_FP32 ? run_batch_fused<float>(ctxs, n, nsteps, nullptr)
                                          : run_batch_fused<double>(ctxs, n, nsteps, nullptr);
    bool graphs = n > 0;
    for (int i = 0; i < n; i++) graphs = graphs && ctxs[i]->use_graph && !ctxs[i]->prof;
    if (graphs) {
        // whole 4-step cycles: each context's captured cycle on its own stream,
        // the contexts' cycles enqueued in turn so the streams overlap
        for (; nsteps >= 4; nsteps -= 4)
            for (int i = 0; i < n; i++) {
                cgfd_ctx *c = ctxs[i];
                HIPCHK(c, hipSetDevice(c->dev));
                int rc = launch_cycle(c);
                if (rc) return rc;
            }
    }
    for (int t = 0; t < nsteps; t++)
        for (int k = 0; k < 4; k++)
            for (int i = 0; i < n; i++) {
                cgfd_ctx *c = ctxs[i];
                HIPCHK(c, hipSetDevice(c->dev));
                int rc = c->prec == CGFD_FP32 ? run_stage<float>(c) : run_stage<double>(c);
                if (rc) return rc;
            }
    return 0;
}

// The fused batch of cgfd_run_batch without allocating: the batch's argument
// images live in the library's fixed per-device array (kStaticBatchMax
// grids), the contexts' events exist since cgfd_create, and the images are
// written from the stack.  Calls that overlap on one device share the array
// and race; a call waits for the previous call's launches before it
// rewrites the array.
int cgfd_run_batch_static(cgfd_ctx *const *ctxs, int n, int nsteps)
{
    if (!ctxs || n < 0 || nsteps < 0) return CGFD_EINVAL;
    for (int i = 0; i < n; i++) {
        cgfd_ctx *c = ctxs[i];
        if (!c) return CGFD_EINVAL;
        if (!c->inited) return fail(c, CGFD_ESTATE, "cgfd_init must run after the last set_* call");
        if (c->stage != 0) return fail(c, CGFD_ESTATE, "cgfd_run_batch_static called in the middle of a step");
        if ((long)c->currT + nsteps > c->d.nt)
            return fail(c, CGFD_ESTATE, "cannot step past nt (the source tables and traces are sized for nt steps)");
    }
    if (n == 0 || nsteps == 0) return 0;
    if (n > kStaticBatchMax)
        return fail(ctxs[0], CGFD_EINVAL,
                    "cgfd_run_batch_static takes at most " + std::to_string(kStaticBatchMax) + " grids, got " +
                        std::to_string(n));
    if (!batch_fusable(ctxs, n))
        return fail(ctxs[0], CGFD_EINVAL,
                    "cgfd_run_batch_static needs a list that cgfd_batch_fusable accepts (2 or more grids on one "
                    "device, same precision, default variant, metric mode and step phase, none profiling)");
    HIPCHK(ctxs[0], hipSetDevice(ctxs[0]->dev));
    ScratchDev *sd = nullptr;
    int rc = scratch_setup(ctxs[0], &sd);
    if (rc) return rc;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return ctxs[0]->prec == CGFD_FP32 ? run_batch_fused<float>(ctxs, n, nsteps, sd)
                                      : run_batch_fused<double>(ctxs, n, nsteps, sd);
}

// Autotuner: time `trial_steps` steps for each candidate z-chunk count
// around the default and keep the fastest; the state is reset by cgfd_init
// after every trial and at the end, so the context is as freshly initialised.
int cgfd_autotune(cgfd_ctx *c, int trial_steps, int *chosen)
{
    if (!c || trial_steps < 1) return CGFD_EINVAL;
    if (!c->inited || c->currT != 0 || c->stage != 0)
        return fail(c, CGFD_ESTATE, "cgfd_autotune needs a freshly initialised context");
    HIPCHK(c, hipSetDevice(c->dev));
    default_tuning(c);
    // one warm-up step plus the timed ones must fit in the nt the run is sized for
    trial_steps = std::min(trial_steps, c->d.nt - 1);
    if (trial_steps < 1) {
        if (chosen) *chosen = c->zchunks;
        c->args_dirty = true;
        return 0;
    }
    const int zc0 = c->zchunks, zmax = std::max(1, c->g.nk / 2);
    std::vector<int> cand = {std::max(1, zc0 / 2), zc0, std::min(zmax, zc0 * 2), std::min(zmax, zc0 * 4)};
    std::sort(cand.begin(), cand.end());
    cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
    hipEvent_t e0, e1;
    HIPCHK(c, hipEventCreate(&e0));
    HIPCHK(c, hipEventCreate(&e1));
    int best = zc0;
    float best_ms = -1;
    int rc = 0;
    // one warm-up step, then trial_steps timed, then the state is reset
    auto trial = [&](float *ms) -> int {
        c->args_dirty = true;
        int r;
        if ((r = cgfd_run_stages(c, 4)) != 0) return r; // warm the instantiations
        if (hipEventRecord(e0, c->stream) != hipSuccess) return fail(c, CGFD_EHIP, "hipEventRecord");
        if ((r = cgfd_run_stages(c, 4 * trial_steps)) != 0) return r;
        if (hipEventRecord(e1, c->stream) != hipSuccess) return fail(c, CGFD_EHIP, "hipEventRecord");
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(ms, e0, e1) != hipSuccess)
            return fail(c, CGFD_EHIP, "event timing failed");
        return cgfd_init(c);
    };
    for (int zc : cand) {
        c->zchunks = zc;
        float ms = 0;
        if ((rc = trial(&ms)) != 0) break;
        if (best_ms < 0 || ms < best_ms) {
            best_ms = ms;
            best = zc;
        }
    }
    c->zchunks = best;
    c->zc_tuned = true;
    c->args_dirty = true;
    // then the LDS kernel's other tile width (FP32) at those chunks, unless
    // cgfd_set_tile_width fixed one
    if (!rc && c->prec == CGFD_FP32 && c->variant == 0 && c->tile_w == 0) {
        c->tile_pref = tile_width(c) == 64 ? 32 : 64;
        float ms = 0;
        rc = trial(&ms);
        if (rc || !(ms < best_ms)) c->tile_pref = 0;
        c->args_dirty = true;
    }
    // then eager stepping against graph replay of 4-step cycles (one warm-up
    // cycle captures the graph), when the run is long enough for both
    if (!rc && trial_steps >= 4 && c->d.nt >= 8) {
        float ms[2] = {0, 0};
        for (int gr = 0; gr < 2 && !rc; gr++) {
            if ((rc = cgfd_init(c)) != 0) break;
            c->use_graph = gr == 1;
            if ((rc = cgfd_run(c, 4)) != 0) break;
            if ((rc = hipEventRecord(e0, c->stream)) != hipSuccess) { rc = fail(c, CGFD_EHIP, "hipEventRecord"); break; }
            if ((rc = cgfd_run(c, 4)) != 0) break;
            if ((rc = hipEventRecord(e1, c->stream)) != hipSuccess) { rc = fail(c, CGFD_EHIP, "hipEventRecord"); break; }
            if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms[gr], e0, e1) != hipSuccess)
                rc = fail(c, CGFD_EHIP, "event timing failed");
        }
        c->use_graph = !rc && ms[1] < ms[0];
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (!rc) rc = cgfd_init(c);
    if (!c->use_graph) drop_graphs(c);
    if (chosen) *chosen = best;
    return rc;
}

int cgfd_sync(cgfd_ctx *c)
{
    if (!c) return CGFD_EINVAL;
    HIPCHK(c, hipSetDevice(c->dev));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int cgfd_current_step(const cgfd_ctx *c) { return c ? c->currT : -1; }
int cgfd_current_stage(const cgfd_ctx *c) { return c ? c->stage : -1; }

// which: 0 = FW, 1 = W (output of the last stage), 2 = tW
static int which_set(const cgfd_ctx *c, int which)
{
    if (which == 0) return SET_FW;
    if (which == 2) return SET_TW;
    if (which == 1) return c->stage == 0 ? SET_FW : stage_out(c->stage - 1);
    return -1;
}

int cgfd_get_field(cgfd_ctx *c, int which, int comp, void *out)
{
    if (!c || !out || comp < 0 || comp >= NF) return fail(c, CGFD_EINVAL, "bad field request");
    const int s = which_set(c, which);
    if (s < 0) return fail(c, CGFD_EINVAL, "which must be 0, 1 or 2");
    HIPCHK(c, hipSetDevice(c->dev));
    return download_field(c, c->field_ptr(s, comp), out);
}

int cgfd_set_field(cgfd_ctx *c, int which, int comp, const void *in)
{
    if (!c || !in || comp < 0 || comp >= NF) return fail(c, CGFD_EINVAL, "bad field request");
    if (which != 0 && which != 2) return fail(c, CGFD_EINVAL, "only FW (0) and tW (2) can be set");
    if (!c->inited) return fail(c, CGFD_ESTATE, "set_field needs an initialised context");
    HIPCHK(c, hipSetDevice(c->dev));
    return upload_field(c, in, c->field_ptr(which_set(c, which), comp));
}

long cgfd_aux_size(const cgfd_ctx *c, int axis)
{
    if (!c || axis < 0 || axis > 2) return -1;
    const long NX = c->g.ni + 6, NY = c->g.nj + 6, NZ = c->g.nk + 6;
    if (axis == 0) return (long)c->npml[0] * NY * NZ;
    if (axis == 1) return (long)c->npml[1] * NX * NZ;
    return (long)c->npml[2] * NX * NY;
}

// index of the engine's auxiliary entry for reference packed entry (p, u, v)
static long aux_engine_index(const cgfd_ctx *c, int axis, int pidx, int a, int b)
{
    // axis 0: (px, y, z); axis 1: (py, x, z); axis 2: (pz, x, y), all 0-based valid
    if (axis == 0) return ((long)b * c->g.nj + a) * c->npml[0] + pidx;
    if (axis == 1) return ((long)b * c->npml[1] + pidx) * c->g.nxp + a;
    return ((long)pidx * c->g.nj + b) * c->g.nxp + a;
}

static int aux_xfer(cgfd_ctx *c, int which, int axis, int comp, void *host, bool get)
{
    if (!c || !host || axis < 0 || axis > 2 || comp < 0 || comp >= NF) return fail(c, CGFD_EINVAL, "bad aux request");
    const int s = which_set(c, which);
    if (s < 0) return fail(c, CGFD_EINVAL, "which must be 0, 1 or 2");
    if (!c->aux[axis]) return 0;
    HIPCHK(c, hipSetDevice(c->dev));
    const long n = c->astr[axis];
    std::vector<char> buf((size_t)n * c->esz);
    char *dev = c->aux_ptr(axis, s) + (size_t)comp * n * c->esz;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (get) HIPCHK(c, hipMemcpy(buf.data(), dev, buf.size(), hipMemcpyDeviceToHost));
    const int NX = c->g.ni + 6, NY = c->g.nj + 6, NZ = c->g.nk + 6;
    const int na = axis == 0 ? c->g.nj : c->g.ni;          // second packed index (valid)
    const int nb = axis == 2 ? c->g.nj : c->g.nk;          // third packed index (valid)
    const int Ma = axis == 0 ? NY : NX, Mb = axis == 2 ? NY : NZ;
    if (get) memset(host, 0, (size_t)cgfd_aux_size(c, axis) * c->esz);
    for (int p = 0; p < c->npml[axis]; p++)
        for (int a = 0; a < na; a++)
            for (int b = 0; b < nb; b++) {
                const long ref = ((long)p * Ma + a + 3) * Mb + b + 3;
                const long eng = aux_engine_index(c, axis, p, a, b);
                if (get)
                    memcpy((char *)host + ref * c->esz, buf.data() + eng * c->esz, c->esz);
                else
                    memcpy(buf.data() + eng * c->esz, (const char *)host + ref * c->esz, c->esz);
            }
    if (!get) HIPCHK(c, hipMemcpy(dev, buf.data(), buf.size(), hipMemcpyHostToDevice));
    return 0;
}

int cgfd_get_aux(cgfd_ctx *c, int which, int axis, int comp, void *out)
{
    return aux_xfer(c, which, axis, comp, out, true);
}

int cgfd_set_aux(cgfd_ctx *c, int which, int axis, int comp, const void *in)
{
    if (c && which != 0 && which != 2) return fail(c, CGFD_EINVAL, "only FA (0) and tA (2) can be set");
    if (c && !c->inited) return fail(c, CGFD_ESTATE, "set_aux needs an initialised context");
    return aux_xfer(c, which, axis, comp, const_cast<void *>(in), false);
}

int cgfd_get_seismo(cgfd_ctx *c, void *out)
{
    if (!c || !out) return CGFD_EINVAL;
    if (!c->nrecv) return 0;
    HIPCHK(c, hipSetDevice(c->dev));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(out, c->seis, (size_t)NF * c->nrecv * c->d.nt * c->esz, hipMemcpyDeviceToHost));
    return 0;
}

int cgfd_get_snap(cgfd_ctx *c, int isnap, int comp, void *out)
{
    if (!c || !out || isnap < 0 || isnap >= (int)c->snaps.size()) return fail(c, CGFD_EINVAL, "bad snapshot index");
    const cgfd_ctx::Snapshot &sn = c->snaps[isnap];
    if (comp < sn.c0 || comp >= sn.c0 + sn.ncomp)
        return fail(c, CGFD_EINVAL, "component " + std::to_string(comp) + " is not recorded by this snapshot");
    HIPCHK(c, hipSetDevice(c->dev));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const size_t slab = (size_t)sn.ncomp * sn.npts;
    std::vector<char> buf(slab * sn.ntime * c->esz);
    HIPCHK(c, hipMemcpy(buf.data(), sn.buf, buf.size(), hipMemcpyDeviceToHost));
    const size_t es = c->esz, ci = (size_t)(comp - sn.c0);
    char *o = (char *)out;
    for (int t = 0; t < sn.ntime; t++)
        for (int p = 0; p < sn.npts; p++)
            memcpy(o + ((size_t)p * sn.ntime + t) * es, buf.data() + ((size_t)t * slab + ci * sn.npts + p) * es, es);
    return 0;
}

} // extern "C"

// seismo.nc as seisplot::point_export writes it at the end of a run
// (seisplot.cpp:672-684 with seisnc.cpp:1336-1365): every step `it` of the
// run goes to record it/tinv with currTime = (it+1)*stept, so with tinv > 1 a
// record keeps the last step written into it.
template <typename R>
static int write_seismo_t(cgfd_ctx *c, const char *path, int tinv, const int *pnum, const int *lnum, const void *posx,
                          const void *posy, const void *posz)
{
    const int T = c->currT, np = c->nrecv, nt = c->d.nt;
    const int nrec = T > 0 ? (T - 1) / tinv + 1 : 0;
    std::vector<R> seis((size_t)NF * np * nt);
    if (np) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        HIPCHK(c, hipMemcpy(seis.data(), c->seis, seis.size() * sizeof(R), hipMemcpyDeviceToHost));
    }
    std::vector<R> wave((size_t)NF * nrec * np), tcur(nrec);
    const R stept = (R)c->d.dt;
    for (int r = 0; r < nrec; r++) {
        const int it = std::min(r * tinv + tinv - 1, T - 1);
        tcur[r] = (R)(it + 1) * stept;
        for (int k = 0; k < NF; k++)
            for (int p = 0; p < np; p++)
                wave[((size_t)k * nrec + r) * np + p] = seis[((size_t)k * np + p) * nt + it];
    }
    std::vector<int> pn(np), ln(np, 1);
    std::vector<R> zero(np, (R)0);
    for (int p = 0; p < np; p++) pn[p] = p;
    const void *w[NF];
    for (int k = 0; k < NF; k++) w[k] = wave.data() + (size_t)k * nrec * np;
    int rc = cgfd_nc_seismo(path, c->prec, nrec, np, w, tcur.data(), pnum ? pnum : pn.data(), lnum ? lnum : ln.data(),
                            posx ? posx : zero.data(), posy ? posy : zero.data(), posz ? posz : zero.data());
    return rc ? fail(c, rc, std::string("cannot write ") + path) : 0;
}

// snap_NNN.nc as seisplot::snap_export writes it (seisplot.cpp:700-760 with
// seisnc.cpp:1448-1490): ceil(steps/tinv) records, record s holding time slot
// s with currTime = (s*tinv+1)*stept.
template <typename R> static int write_snap_t(cgfd_ctx *c, int isnap, const char *path)
{
    const cgfd_ctx::Snapshot &sn = c->snaps[isnap];
    const int tinv = sn.spec[9];
    const int nrec = std::min((c->currT + tinv - 1) / tinv, sn.ntime);
    const size_t slab = (size_t)sn.ncomp * sn.npts;
    std::vector<R> buf(slab * std::max(nrec, 1));
    if (nrec) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        HIPCHK(c, hipMemcpy(buf.data(), sn.buf, slab * nrec * sizeof(R), hipMemcpyDeviceToHost));
    }
    std::vector<R> wave((size_t)NF * nrec * sn.npts), tcur(nrec);
    const R stept = (R)c->d.dt;
    for (int r = 0; r < nrec; r++) {
        tcur[r] = (R)(r * tinv + 1) * stept;
        for (int ci = 0; ci < sn.ncomp; ci++)
            memcpy(wave.data() + ((size_t)(sn.c0 + ci) * nrec + r) * sn.npts, buf.data() + (size_t)r * slab + (size_t)ci * sn.npts,
                   sn.npts * sizeof(R));
    }
    const void *w[NF];
    for (int k = 0; k < NF; k++) w[k] = wave.data() + (size_t)k * nrec * sn.npts;
    int rc = cgfd_nc_snap(path, c->prec, sn.spec, nrec, w, tcur.data());
    return rc ? fail(c, rc, std::string("cannot write ") + path) : 0;
}

extern "C" {

int cgfd_write_seismo(cgfd_ctx *c, const char *path, int tinv, const int *pnum, const int *lnum, const void *posx,
                      const void *posy, const void *posz)
{
    if (!c || !path || tinv < 1) return fail(c, CGFD_EINVAL, "bad seismogram output arguments");
    if (!c->nrecv) return fail(c, CGFD_ESTATE, "no receivers are set");
    HIPCHK(c, hipSetDevice(c->dev));
    return c->prec == CGFD_FP32 ? write_seismo_t<float>(c, path, tinv, pnum, lnum, posx, posy, posz)
                                : write_seismo_t<double>(c, path, tinv, pnum, lnum, posx, posy, posz);
}

int cgfd_write_snap(cgfd_ctx *c, int isnap, const char *path)
{
    if (!c || !path || isnap < 0 || isnap >= (int)c->snaps.size()) return fail(c, CGFD_EINVAL, "bad snapshot index");
    HIPCHK(c, hipSetDevice(c->dev));
    return c->prec == CGFD_FP32 ? write_snap_t<float>(c, isnap, path) : write_snap_t<double>(c, isnap, path);
}

int cgfd_write_peakvel(cgfd_ctx *c, const char *path, int vz_from_vy)
{
    if (!c || !path) return fail(c, CGFD_EINVAL, "bad peak-velocity output arguments");
    const long NX = c->g.ni + 6, NY = c->g.nj + 6;
    std::vector<char> pv((size_t)3 * NX * NY * c->esz), out((size_t)3 * c->g.ni * c->g.nj * c->esz);
    int rc = cgfd_get_pgv(c, pv.data());
    if (rc) return rc;
    const size_t es = c->esz;
    for (int k = 0; k < 3; k++)
        for (int x = 0; x < c->g.ni; x++)
            for (int y = 0; y < c->g.nj; y++)
                memcpy(out.data() + (((size_t)k * c->g.ni + x) * c->g.nj + y) * es,
                       pv.data() + (((size_t)k * NX + x + 3) * NY + y + 3) * es, es);
    const size_t plane = (size_t)c->g.ni * c->g.nj * es;
    // PV_export (seisnc.cpp:1703-1706) fills Vz from the Vy map
    rc = cgfd_nc_peakvel(path, c->prec, c->g.ni, c->g.nj, out.data(), out.data() + plane,
                         out.data() + (vz_from_vy ? 1 : 2) * plane);
    return rc ? fail(c, rc, std::string("cannot write ") + path) : 0;
}

// maps stored per valid (x,y) as [slot][nj][nxp] -> reference [slot][NX][NY]
static int get_xy_maps(cgfd_ctx *c, const char *dev, int nslots, int slot_minor, void *out)
{
    HIPCHK(c, hipSetDevice(c->dev));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::vector<char> buf((size_t)nslots * c->cstr * c->esz);
    HIPCHK(c, hipMemcpy(buf.data(), dev, buf.size(), hipMemcpyDeviceToHost));
    const long NX = c->g.ni + 6, NY = c->g.nj + 6;
    memset(out, 0, (size_t)nslots * NX * NY * c->esz);
    for (int s = 0; s < nslots; s++)
        for (int x = 0; x < c->g.ni; x++)
            for (int y = 0; y < c->g.nj; y++) {
                const long src = (long)s * c->cstr + (long)y * c->g.nxp + x;
                const long xy = (long)(x + 3) * NY + (y + 3);
                const long dst = slot_minor ? xy * nslots + s : (long)s * NX * NY + xy;
                memcpy((char *)out + dst * c->esz, buf.data() + src * c->esz, c->esz);
            }
    return 0;
}

int cgfd_get_pgv(cgfd_ctx *c, void *out)
{
    if (!c || !out) return CGFD_EINVAL;
    return get_xy_maps(c, c->pgv, 3, 0, out);
}

int cgfd_get_cov(cgfd_ctx *c, void *covx, void *covy)
{
    if (!c || !covx || !covy) return CGFD_EINVAL;
    int rc = get_xy_maps(c, c->cov, 9, 1, covx);
    if (rc) return rc;
    return get_xy_maps(c, c->cov + (size_t)9 * c->cstr * c->esz, 9, 1, covy);
}

int cgfd_profile(cgfd_ctx *c, int on)
{
    if (!c) return CGFD_EINVAL;
    HIPCHK(c, hipSetDevice(c->dev));
    fold_events(c);
    c->prof = on != 0;
    c->ms_stage = c->ms_timg = 0;
    c->n_stage = c->n_timg = 0;
    return 0;
}

int cgfd_profile_read(cgfd_ctx *c, double *ms_stage, long *n_stage, double *ms_timg, long *n_timg)
{
    if (!c) return CGFD_EINVAL;
    HIPCHK(c, hipSetDevice(c->dev));
    fold_events(c);
    if (ms_stage) *ms_stage = c->ms_stage;
    if (n_stage) *n_stage = c->n_stage;
    if (ms_timg) *ms_timg = c->ms_timg;
    if (n_timg) *n_timg = c->n_timg;
    return 0;
}

int cgfd_set_tile_width(cgfd_ctx *c, int width)
{
    if (!c) return CGFD_EINVAL;
    if (width != 0 && width != 32 && width != 64) return fail(c, CGFD_EINVAL, "tile width must be 0, 32 or 64");
    if (width == 32 && c->prec == CGFD_FP64) return fail(c, CGFD_EINVAL, "the FP64 kernels run 64-column tiles");
    if (width != c->tile_w) {
        c->tile_w = width;
        c->args_dirty = true; // the next run rewrites the image and drops captured cycles
    }
    return 0;
}

int cgfd_tile_width(const cgfd_ctx *c) { return c ? tile_width(c) : CGFD_EINVAL; }

int cgfd_set_tuning(cgfd_ctx *c, int zchunks, int rows)
{
    if (!c) return CGFD_EINVAL;
    if (rows != 0 && rows != 1 && rows != 2 && rows != 4) return fail(c, CGFD_EINVAL, "rows must be 1, 2 or 4");
    if (zchunks < 0 || zchunks > c->g.nk) return fail(c, CGFD_EINVAL, "zchunks out of range");
    if (rows) c->rows = rows;
    if (zchunks) {
        c->zchunks = zchunks;
        c->zc_tuned = true;
    }
    c->args_dirty = true;
    return 0;
}

int cgfd_set_variant(cgfd_ctx *c, int variant)
{
    if (!c) return CGFD_EINVAL;
    if (variant < 0 || variant > 2)
        return fail(c, CGFD_EINVAL, "variant must be 0 (LDS), 1 (direct) or 2 (split velocity/stress)");
    if (variant != c->variant) {
        c->variant = variant;
        // captured cycles launch the old variant's kernels: the next run
        // rewrites the argument image, which drops them after a stream sync
        c->args_dirty = true;
    }
    return 0;
}

double cgfd_stage_bytes(const cgfd_ctx *c)
{
    if (!c) return -1;
    const double nv = (double)c->g.ni * c->g.nj * c->g.nk;
    const double npml = (double)c->npml[0] * c->g.nj * c->g.nk + (double)c->npml[1] * c->g.ni * c->g.nk +
                        (double)c->npml[2] * c->g.ni * c->g.nj;
    const double scale = c->esz / 4.0;
    // the constant-horizontal-metrics kernel skips 6 of the 9 metric reads
    const double cell = (c->hmet && c->variant == 0 && c->use_hmet) ? 192.0 - 24.0 : 192.0;
    return scale * (cell * nv + 144.0 * npml);
}

int cgfd_metrics_specialized(const cgfd_ctx *c) { return c ? (c->hmet && c->use_hmet && c->variant == 0) : 0; }

int cgfd_set_specialize(cgfd_ctx *c, int on)
{
    if (!c) return CGFD_EINVAL;
    c->use_hmet = on != 0;
    c->args_dirty = true;
    return 0;
}

void *cgfd_stream(const cgfd_ctx *c) { return c ? (void *)c->stream : nullptr; }

} // extern "C"
