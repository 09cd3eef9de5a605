// cgl_api.hip -- the C ABI of libcgl (declared in include/cgl.h).
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <mutex>
#include <new>
#include <vector>

#include "cgl_impl.hpp"

namespace {

thread_local char g_err[512] = "";

int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

#define HIPCHK(expr)                                                                                             \
    do {                                                                                                         \
        hipError_t e_ = (expr);                                                                                  \
        if (e_ != hipSuccess) return fail(CGL_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));          \
    } while (0)

// Dispatch on the storage precision.
template <class Fn> void dispatch(cgl_handle *h, Fn fn)
{
    switch (h->p.precision) {
    case CGL_PREC_F64: fn(cgl::ImplF64{}); break;
    case CGL_PREC_F32: fn(cgl::ImplF32{}); break;
    default: fn(cgl::ImplF16{}); break;
    }
}

int check_launch(const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(CGL_ERR_HIP, "%s launch failed: %s", what, hipGetErrorString(e));
    return CGL_OK;
}

// Device allocation of the handle's buffers.  CGL_POISON=<mask> in the
// environment fills each new buffer of the classes in the mask with 0xff
// bytes (NaN in FP16, FP32 and FP64, -1 in integers) before use, so that a
// kernel which reads memory no step has written shows it in its results
// instead of reading whatever the allocator hands back (zero pages in a
// fresh process).  Classes: 1 populations, 2 phase field, 4 factorised
// state, 8 particle fields and masks, 16 particle state and partial sums,
// 32 scratch of downloads and diagnostics, 64 scatter-form sums, 128 tables.
enum : unsigned { PZ_POPS = 1, PZ_PHI = 2, PZ_FS = 4, PZ_PFIELD = 8, PZ_PART = 16, PZ_SCR = 32, PZ_SC = 64, PZ_TAB = 128 };
hipError_t dmalloc(void **p, size_t bytes, unsigned cls, hipStream_t st)
{
    static const unsigned poison = [] {
        const char *env = getenv("CGL_POISON");
        return env ? (unsigned)strtoul(env, nullptr, 0) : 0u;
    }();
    hipError_t e = hipMalloc(p, bytes);
    if (e == hipSuccess && (poison & cls)) e = hipMemsetAsync(*p, 0xff, bytes, st);
    return e;
}

// Scratch for host-layout conversions, sized for the largest box a call can
// address (scratch_n nodes).
int ensure_scratch(cgl_handle *h, bool mac, bool aos, bool iflag)
{
    const size_t n = h->scratch_n;
    if (mac && !h->mac) HIPCHK(dmalloc((void **)&h->mac, 7 * n * sizeof(double), PZ_SCR, h->stream));
    if (aos && !h->aos) HIPCHK(dmalloc((void **)&h->aos, (size_t)cgl::Q * n * sizeof(double), PZ_SCR, h->stream));
    if (iflag && !h->iflag) HIPCHK(dmalloc((void **)&h->iflag, n * sizeof(int), PZ_SCR, h->stream));
    return CGL_OK;
}

// Bytes per node of the largest array the kernels address through a 32-bit
// byte offset from one base: a group of four direction arrays (grp/grp_off),
// and in the factorised layout the per-node recolour scalars (16 B in
// FP32/FP16, 48 B in FP64).  The largest offset, bytes x N minus one element,
// fits 32 bits when bytes x N <= 4 GiB.
double addressed_bytes(const cgl_params *p)
{
    const double es = p->precision == CGL_PREC_F64 ? 8.0 : (p->precision == CGL_PREC_F32 ? 4.0 : 2.0);
    double b = 4.0 * es;
    if (p->layout == CGL_LAYOUT_FS) {
        const double ns = p->precision == CGL_PREC_F64 ? 48.0 : 16.0;
        if (ns > b) b = ns;
    }
    return b;
}

int set_device(cgl_handle *h)
{
    HIPCHK(hipSetDevice(h->p.device));
    return CGL_OK;
}

// One time step, then swap buffers.  Gather form: K1 then K2.  Scatter form
// (FP32/FP16 factorised boxes, once the state is factorised; with particles
// once their start-up step is done): the step kernels and the phase-field
// pass; `last` = the run's last step, which writes the recolour scalars so
// that the state can be read back.
void enqueue_step(cgl_handle *h, bool last = true)
{
    dispatch(h, [&](auto impl) {
        if (h->sc && h->state_layout == cgl::FS && !(h->npar && (!h->primed || h->cold))) {
            if (h->ext)
                decltype(impl)::step_sc_ext(h, last);
            else
                decltype(impl)::step_sc(h, last);
        } else if (h->nbox) {
            decltype(impl)::batch_moments(h);
            decltype(impl)::batch_collide(h);
            h->phi_ok = 0;
        } else if (h->ext) {
            decltype(impl)::step_ext(h);
        } else {
            decltype(impl)::moments(h, false);
            decltype(impl)::collide(h);
            h->phi_ok = 0;
        }
    });
    h->cur ^= 1;
    h->phicur ^= 1;
    h->steps_done++;
}

// Steps that must run eagerly before steady graph replays: the layout
// conversion, the particles' start-up step, and the scatter form's first
// step (its phase field comes from a moments pass).
bool needs_eager(const cgl_handle *h)
{
    return h->state_layout != h->layout || (h->npar && (!h->primed || h->cold)) || (h->sc && !h->phi_ok);
}

// Captured steps hold their kernels' by-value arguments (particle scalars,
// geometry): drop the graph when those change.
int drop_graph(cgl_handle *h)
{
    if (h->gexec) {
        HIPCHK(hipGraphExecDestroy(h->gexec));
        h->gexec = nullptr;
        h->gkey = -1;
    }
    return CGL_OK;
}

// Entry points that read the state back need its recolour scalars.
// The scatter form's sums are 32-bit fixed point at 2^29 units per rho0: a
// colour density of 4 rho0 or more wraps, and the finishing pass raises the
// handle's range flag.  Checked wherever the host reads or waits on the
// state; cleared when a new state is installed.
int check_range(const cgl_handle *h)
{
    if (!h->sc || !h->sb.range) return CGL_OK;
    int v = 0;
    HIPCHK(hipMemcpyAsync(&v, h->sb.range, sizeof v, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    if (v)
        return fail(CGL_ERR_RANGE, "a colour density reached 4 rho0, beyond the fixed-point sums of the scatter "
                                   "form; the state is lost (run with CGL_SCATTER=0 for such densities)");
    return CGL_OK;
}

// (Callers that install a state ignore its result: a HIP error here shows
// again at the next call.)
int clear_range(cgl_handle *h)
{
    if (h->sb.range) HIPCHK(hipMemsetAsync(h->sb.range, 0, sizeof(int), h->stream));
    return CGL_OK;
}

// A staged step between its stages holds moments, masks and particle state
// of different time levels: only the particle, mask and band read-backs may
// look at it.
int no_open_stage(const cgl_handle *h)
{
    if (h->stage > STAGE_CALMACROS)
        return fail(CGL_ERR_ARG, "a staged step is open: finish it with cgl_collide_stream");
    return CGL_OK;
}

int state_readable(const cgl_handle *h)
{
    if (h->scal_stale) return fail(CGL_ERR_HIP, "internal: the state's recolour scalars were not written");
    if (int rc = no_open_stage(h)) return rc;
    return check_range(h);
}

// Fixed-point partial sums of the scatter step, for FP32/FP16 factorised
// boxes (single, or every box of a batch) whose workgroup boxes have at
// least two rows and two lanes; with wall faces and particles for single
// boxes without mirror faces, at the standard tile height.
int setup_scatter(cgl_handle *h)
{
    const char *env = getenv("CGL_SCATTER"); // 0: gather form (moments pass) on every step
    const bool want = env ? atoi(env) != 0 : true;
    if (!want || h->p.precision == CGL_PREC_F64 || h->p.layout != CGL_LAYOUT_FS) return CGL_OK;
    if (h->ext) {
        for (int a = 0; a < 3; ++a)
            if (h->E.mirror[a][0] || h->E.mirror[a][1]) return CGL_OK;
        if (h->nbox || (h->tile_rows && h->tile_rows != cgl::KTY<float>)) return CGL_OK;
        h->ky = cgl::KTY<float>;
    }
    if (h->nbox) {
        for (int b = 0; b < h->nbox; ++b)
            if (!cgl::sc_shape_ok(h->hboxes[b].L.ny, h->hboxes[b].L.nz)) return CGL_OK;
    } else if (!cgl::sc_shape_ok(h->p.ny, h->p.nz, h->ky)) {
        return CGL_OK;
    }
    int2 **bufs[cgl::SC_NBUF] = {&h->sb.own,   &h->sb.fxlo,  &h->sb.fxhi,  &h->sb.fy,    &h->sb.fz,
                                 &h->sb.exylo, &h->sb.exyhi, &h->sb.exzlo, &h->sb.exzhi, &h->sb.eyz};
    for (int k = 0; k < cgl::SC_NBUF; ++k) HIPCHK(dmalloc((void **)bufs[k], h->N * sizeof(int2), PZ_SC, h->stream));
    HIPCHK(dmalloc((void **)&h->sb.range, 256, PZ_SC, h->stream));
    if (h->npar) HIPCHK(dmalloc((void **)&h->bnrm, 3 * h->N * sizeof(float), PZ_SC, h->stream));
    h->sc = 1;
    return clear_range(h);
}

// Streaming copy, 16 B per thread per iteration: the reference kernel for the
// box's achievable HBM bandwidth and for calibrating PMC byte counters.
// One float4 per thread over the whole buffer: every wave has its load in
// flight at once, which streams faster on this device (about 6.2 TB/s) than
// a grid-stride loop over a few workgroups per CU (about 4.5 TB/s).
__global__ __launch_bounds__(256) void k_copy(const float4 *__restrict__ a, float4 *__restrict__ b, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

} // namespace

extern "C" {

int cgl_probe_copy_bandwidth(int32_t device, int64_t bytes, int32_t reps, double *gbps)
{
    if (bytes < (1 << 20) || reps < 1 || !gbps) return fail(CGL_ERR_ARG, "bad arguments");
    HIPCHK(hipSetDevice(device));
    const size_t n = (size_t)bytes / sizeof(float4);
    float4 *a = nullptr, *b = nullptr;
    HIPCHK(hipMalloc((void **)&a, n * sizeof(float4)));
    hipError_t e = hipMalloc((void **)&b, n * sizeof(float4));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipMemset(a, 0, n * sizeof(float4));
    float ms = 0.f;
    if (e == hipSuccess) {
        const unsigned blocks = (unsigned)((n + 255) / 256);
        k_copy<<<blocks, 256>>>(a, b, n); // warm-up
        e = hipEventRecord(e0, 0);
        for (int r = 0; r < reps && e == hipSuccess; ++r) k_copy<<<blocks, 256>>>(a, b, n);
        if (e == hipSuccess) e = hipGetLastError();
        if (e == hipSuccess) e = hipEventRecord(e1, 0);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (e != hipSuccess) return fail(CGL_ERR_HIP, "copy probe failed: %s", hipGetErrorString(e));
    *gbps = 2.0 * (double)(n * sizeof(float4)) * reps / (ms * 1e-3) / 1e9;
    return CGL_OK;
}


const char *cgl_last_error(void) { return g_err; }

static int alloc_handle(cgl_t **out, const cgl_params *p, size_t N, int walls);

int cgl_create(cgl_t **out, const cgl_params *p)
{
    if (!out || !p) return fail(CGL_ERR_ARG, "null argument");
    *out = nullptr;
    if (p->nx < 1 || p->ny < 1 || p->nz < 1) return fail(CGL_ERR_ARG, "box must be at least 1x1x1");
    double nn = (double)p->nx * p->ny * p->nz;
    if (nn > 2.0e9) return fail(CGL_ERR_ARG, "box of %.3g nodes exceeds the 32-bit node index", nn);
    if (nn * addressed_bytes(p) > 4294967296.0)
        return fail(CGL_ERR_UNSUPPORTED, "box of %.3g nodes exceeds the 4 GiB per-array addressing", nn);
    if (p->precision < CGL_PREC_F64 || p->precision > CGL_PREC_F16)
        return fail(CGL_ERR_ARG, "unknown precision %d", p->precision);
    int walls = 0, mirrors = 0;
    const int nn3[3] = {p->nx, p->ny, p->nz};
    for (int a = 0; a < 3; ++a) {
        for (int s = 0; s < 2; ++s) {
            const int f = p->bc[2 * a + s];
            if (f != CGL_BC_PERIODIC && f != CGL_BC_WALL && f != CGL_BC_MIRROR)
                return fail(CGL_ERR_ARG, "face %d: unknown type %d", 2 * a + s, f);
            walls |= f == CGL_BC_WALL;
            mirrors |= f == CGL_BC_MIRROR;
        }
        if ((p->bc[2 * a] == CGL_BC_PERIODIC) != (p->bc[2 * a + 1] == CGL_BC_PERIODIC))
            return fail(CGL_ERR_ARG, "axis %d: periodic faces come in pairs", a);
        if (p->bc[2 * a] != CGL_BC_PERIODIC && nn3[a] < 4)
            return fail(CGL_ERR_ARG, "axis %d: a bounded axis needs two ghost layers and two live nodes", a);
    }
    if (p->npar < 0) return fail(CGL_ERR_ARG, "npar must be >= 0");
    if (!(p->rho0 > 0) || !(p->nu_r > 0) || !(p->nu_b > 0)) return fail(CGL_ERR_ARG, "rho0, nu_r, nu_b must be > 0");
    if (p->layout != CGL_LAYOUT_RB && p->layout != CGL_LAYOUT_FS) return fail(CGL_ERR_ARG, "unknown layout %d", p->layout);
    if (p->tile_rows != 0 && p->tile_rows != 4 && p->tile_rows != 8 && p->tile_rows != 16)
        return fail(CGL_ERR_ARG, "tile_rows must be 0, 4, 8 or 16");
    int rc = alloc_handle(out, p, (size_t)p->nx * p->ny * p->nz, walls | mirrors);
    if (rc) return rc;
    if ((rc = setup_scatter(*out))) {
        cgl_destroy(*out);
        *out = nullptr;
    }
    return rc;
}

// Allocation of a handle's device state for N nodes (one box, or a packed
// batch of boxes).
static int alloc_handle(cgl_t **out, const cgl_params *p, size_t N, int walls)
{
    cgl_handle *h = new (std::nothrow) cgl_handle();
    if (!h) return fail(CGL_ERR_ARG, "out of host memory");
    memset(h, 0, sizeof *h);
    h->p = *p;
    h->L.nx = p->nx;
    h->L.ny = p->ny;
    h->L.nz = p->nz;
    h->N = N;
    h->scratch_n = N;
    h->gkey = -1;
    {
        const char *env = getenv("CGL_GRAPHS"); // 0: launch every step's kernels directly
        h->use_graph = env ? atoi(env) != 0 : 1;
    }
    h->L.N = (long long)p->nx * p->ny * p->nz;
    h->L.rho0 = p->rho0;
    h->L.sigma = p->sigma;
    h->L.beta = p->beta;
    h->L.nu_r = p->nu_r;
    h->L.nu_b = p->nu_b;
    {
        const char *env = getenv("CGL_XCD_REMAP");
        h->L.xcd_remap = env ? atoi(env) : 1;
        h->remap_env = env != nullptr;
        h->sc_remap = env ? h->L.xcd_remap : 2;
    }
    h->tile_rows = p->tile_rows;
    h->ky = p->tile_rows ? p->tile_rows : 8;
    h->tune_best_ms = -1.f;
    h->esize = p->precision == CGL_PREC_F64 ? 8 : (p->precision == CGL_PREC_F32 ? 4 : 2);
    h->phisize = p->precision == CGL_PREC_F64 ? 8 : 4;
    h->xc = p->xchunk > 0 ? p->xchunk : 8;
    h->tuning = p->xchunk > 0 ? 0 : 1;
    h->layout = p->layout;
    h->state_layout = cgl::RB;
    h->npar = p->npar;
    h->ext = walls || p->npar > 0;
    for (int a = 0; a < 3; ++a)
        for (int s = 0; s < 2; ++s) {
            h->E.wall[a][s] = p->bc[2 * a + s] == CGL_BC_WALL;
            h->E.mirror[a][s] = p->bc[2 * a + s] == CGL_BC_MIRROR;
        }
    h->E.npar = p->npar;
    h->P.npar = p->npar;
    int rc = set_device(h);
    if (rc) {
        delete h;
        return rc;
    }
    hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    for (int k = 0; k < 4 && e == hipSuccess; ++k) e = hipEventCreate(&h->ev[k]);
    // RB: both population buffers.  FS: one RB buffer (staging of uploaded or
    // initial states) plus the factorised double buffer; with walls or
    // particles both RB buffers, which then hold the red/blue side arrays of
    // the factorised states (the slots of solid nodes).
    // Shifted storage keeps one FP32 colour fraction per slot there (the red
    // array, at least 4 B per value); FP64 red and blue values.
    const int nrb = p->layout == CGL_LAYOUT_RB || h->ext ? 2 : 1;
    const bool fsx = p->layout == CGL_LAYOUT_FS && h->ext, fsx_frac = fsx && p->precision != CGL_PREC_F64;
    for (int b = 0; b < nrb && e == hipSuccess; ++b)
        for (int c = 0; c < 2 && e == hipSuccess; ++c) {
            if (fsx_frac && b == 1 && c == 1) continue; // no blue side array
            const size_t es = fsx_frac && c == 0 && h->esize < 4 ? 4 : h->esize;
            e = dmalloc(&h->pops[b][c], (size_t)cgl::Q * h->N * es, PZ_POPS, h->stream);
        }
    if (p->layout == CGL_LAYOUT_FS) {
        const size_t ns = p->precision == CGL_PREC_F64 ? 6 * sizeof(double) : 4 * sizeof(float);
        for (int b = 0; b < 2 && e == hipSuccess; ++b) e = dmalloc(&h->fsf[b], (size_t)cgl::Q * h->N * h->esize, PZ_FS, h->stream);
        for (int b = 0; b < 2 && e == hipSuccess; ++b) e = dmalloc(&h->fss[b], h->N * ns, PZ_FS, h->stream);
    }
    for (int b = 0; b < 2 && e == hipSuccess; ++b) e = dmalloc(&h->phi[b], h->N * h->phisize, PZ_PHI, h->stream);
    if (h->npar > 0) {
        // masks, NEAR-band macroscopic fields and gradients, particle state
        for (int b = 0; b < 2 && e == hipSuccess; ++b) e = dmalloc((void **)&h->ntype[b], h->N * sizeof(unsigned), PZ_PFIELD, h->stream);
        for (int b = 0; b < 2 && e == hipSuccess; ++b) e = hipMemsetAsync(h->ntype[b], 0, h->N * sizeof(unsigned), h->stream);
        int **ints[3] = {&h->walls_i, &h->pin_i, &h->pout_i};
        for (int k = 0; k < 3 && e == hipSuccess; ++k) e = dmalloc((void **)ints[k], h->N * sizeof(int), PZ_PFIELD, h->stream);
        for (int k = 0; k < 3 && e == hipSuccess; ++k) e = hipMemsetAsync(*ints[k], 0, h->N * sizeof(int), h->stream);
        if (e == hipSuccess) e = dmalloc((void **)&h->E.mac, 6 * h->N * sizeof(double), PZ_PFIELD, h->stream);
        if (e == hipSuccess) e = hipMemsetAsync(h->E.mac, 0, 6 * h->N * sizeof(double), h->stream);
        if (e == hipSuccess) e = dmalloc((void **)&h->E.der, 6 * h->N * sizeof(double), PZ_PFIELD, h->stream);
        if (e == hipSuccess) e = hipMemsetAsync(h->E.der, 0, 6 * h->N * sizeof(double), h->stream);
        const size_t n3 = 3 * (size_t)h->npar * sizeof(double), n1 = (size_t)h->npar * sizeof(double);
        double **v3[11] = {&h->P.center, &h->P.center_old, &h->P.center_prev, &h->P.up, &h->P.rp,
                           &h->P.fp, &h->P.tp, &h->P.frep, &h->P.ten, &h->P.tent, &h->P.center0};
        double **v1[5] = {&h->P.rad, &h->P.mass, &h->P.inertia, &h->P.cita, &h->P.shb};
        for (int k = 0; k < 11 && e == hipSuccess; ++k) e = dmalloc((void **)v3[k], n3, PZ_PART, h->stream);
        for (int k = 0; k < 11 && e == hipSuccess; ++k) e = hipMemsetAsync(*v3[k], 0, n3, h->stream);
        for (int k = 0; k < 5 && e == hipSuccess; ++k) e = dmalloc((void **)v1[k], n1, PZ_PART, h->stream);
        for (int k = 0; k < 5 && e == hipSuccess; ++k) e = hipMemsetAsync(*v1[k], 0, n1, h->stream);
        if (e == hipSuccess) e = dmalloc((void **)&h->P.part_lag, (size_t)h->npar * cgl::LAG_BLOCKS * 7 * sizeof(double), PZ_PART, h->stream);
        if (e == hipSuccess) e = dmalloc((void **)&h->P.part_mom, (size_t)h->npar * cgl::LSLICES * 6 * sizeof(double), PZ_PART, h->stream);
        if (e == hipSuccess) e = dmalloc((void **)&h->P.lcount, 2 * (size_t)h->npar * cgl::LSLICES * sizeof(int), PZ_PART, h->stream);
        h->E.pout = h->pout_i;
        h->E.center = h->P.center;
        h->E.cita = h->P.cita;
    }
    if (e == hipSuccess) e = hipMemsetAsync(h->phi[0], 0, h->N * h->phisize, h->stream);
    if (e == hipSuccess) e = hipMemsetAsync(h->phi[1], 0, h->N * h->phisize, h->stream);
    // every device operation of a handle runs on its stream, which does not
    // order itself with the null stream (hipStreamNonBlocking): the clears
    // above must be complete before the caller's first entry point
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) {
        rc = fail(CGL_ERR_HIP, "device allocation failed: %s", hipGetErrorString(e));
        cgl_destroy(h);
        return rc;
    }
    *out = h;
    return CGL_OK;
}

int cgl_destroy(cgl_t *h)
{
    if (!h) return CGL_OK;
    (void)hipSetDevice(h->p.device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (int b = 0; b < 2; ++b) {
        for (int c = 0; c < 2; ++c)
            if (h->pops[b][c]) (void)hipFree(h->pops[b][c]);
        if (h->phi[b]) (void)hipFree(h->phi[b]);
        if (h->fsf[b]) (void)hipFree(h->fsf[b]);
        if (h->fss[b]) (void)hipFree(h->fss[b]);
    }
    if (h->mac) (void)hipFree(h->mac);
    if (h->npar > 0) {
        void *ptrs[] = {h->ntype[0], h->ntype[1], h->walls_i, h->pin_i, h->pout_i, h->E.mac, h->E.der,
                        h->P.center, h->P.center_old, h->P.center_prev, h->P.up, h->P.rp, h->P.fp, h->P.tp,
                        h->P.frep, h->P.ten, h->P.tent, h->P.rad, h->P.mass, h->P.inertia, h->P.cita, h->P.shb,
                        h->P.part_lag, h->P.part_mom, h->P.center0, h->P.lout, h->P.lmask, h->P.lin, h->P.lcount};
        for (void *q : ptrs)
            if (q) (void)hipFree(q);
    }
    if (h->aos) (void)hipFree(h->aos);
    if (h->iflag) (void)hipFree(h->iflag);
    if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
    if (h->dboxes) (void)hipFree(h->dboxes);
    delete[] h->hboxes;
    int2 *sbufs[cgl::SC_NBUF] = {h->sb.own,   h->sb.fxlo,  h->sb.fxhi,  h->sb.fy,    h->sb.fz,
                                 h->sb.exylo, h->sb.exyhi, h->sb.exzlo, h->sb.exzhi, h->sb.eyz};
    for (int2 *q : sbufs)
        if (q) (void)hipFree(q);
    if (h->sb.range) (void)hipFree(h->sb.range);
    if (h->bnrm) (void)hipFree(h->bnrm);
    for (int k = 0; k < 4; ++k)
        if (h->ev[k]) (void)hipEventDestroy(h->ev[k]);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return CGL_OK;
}

int cgl_upload_state(cgl_t *h, const double *ri, const double *bi)
{
    if (h && h->nbox) return fail(CGL_ERR_ARG, "a batch handle takes the cgl_box_* entry points");
    if (!h || !ri || !bi) return fail(CGL_ERR_ARG, "null argument");
    int rc = set_device(h);
    if (rc) return rc;
    if ((rc = ensure_scratch(h, false, true, false))) return rc;
    const size_t bytes = (size_t)cgl::Q * h->N * sizeof(double);
    // the shift reference of the uploaded state: phi of each node
    double *tmp = nullptr;
    HIPCHK(dmalloc((void **)&tmp, bytes, PZ_SCR, h->stream));
    hipError_t e = hipMemcpyAsync(h->aos, ri, bytes, hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(tmp, bi, bytes, hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) {
        h->cur = 0; // uploaded states are held in RB form in buffer 0
        h->state_layout = cgl::RB;
        h->primed = 0;
        h->stage = STAGE_NONE;
        h->phi_ok = 0;
        h->scal_stale = 0;
        (void)clear_range(h);
        dispatch(h, [&](auto impl) {
            decltype(impl)::phi_aos(h, h->aos, tmp);
            decltype(impl)::pack(h, h->aos, 1);
            decltype(impl)::pack(h, tmp, 0);
        });
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(tmp);
    if (e != hipSuccess) return fail(CGL_ERR_HIP, "upload failed: %s", hipGetErrorString(e));
    return CGL_OK;
}

int cgl_download_state(cgl_t *h, double *ri, double *bi)
{
    if (h && h->nbox) return fail(CGL_ERR_ARG, "a batch handle takes the cgl_box_* entry points");
    if (!h || !ri || !bi) return fail(CGL_ERR_ARG, "null argument");
    int rc = set_device(h);
    if (rc) return rc;
    if ((rc = state_readable(h))) return rc;
    if ((rc = ensure_scratch(h, false, true, false))) return rc;
    const size_t bytes = (size_t)cgl::Q * h->N * sizeof(double);
    for (int red = 1; red >= 0; --red) {
        dispatch(h, [&](auto impl) { decltype(impl)::unpack(h, h->aos, red); });
        if ((rc = check_launch("unpack"))) return rc;
        HIPCHK(hipMemcpyAsync(red ? ri : bi, h->aos, bytes, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
    }
    return CGL_OK;
}

int cgl_init_cylinder(cgl_t *h, double xc, double yc, double r0, double kprof, int32_t nmodes, const double *amp)
{
    if (h && h->nbox) return fail(CGL_ERR_ARG, "a batch handle takes the cgl_box_* entry points");
    if (!h) return fail(CGL_ERR_ARG, "null handle");
    if (nmodes < 0 || nmodes > 16 || (nmodes > 0 && !amp)) return fail(CGL_ERR_ARG, "nmodes must be 0..16");
    int rc = set_device(h);
    if (rc) return rc;
    cgl::CylIC ic{};
    ic.xc = xc;
    ic.yc = yc;
    ic.r0 = r0;
    ic.kprof = kprof;
    ic.nmodes = nmodes;
    for (int m = 0; m < nmodes; ++m) ic.amp[m] = amp[m];
    h->cur = 0; // initial states are held in RB form in buffer 0
    h->state_layout = cgl::RB;
    h->primed = 0;
    h->stage = STAGE_NONE;
    h->phi_ok = 0;
    h->scal_stale = 0;
    (void)clear_range(h);
    dispatch(h, [&](auto impl) { decltype(impl)::init_cylinder(h, ic); });
    if ((rc = check_launch("init_cylinder"))) return rc;
    HIPCHK(hipStreamSynchronize(h->stream));
    return CGL_OK;
}

int cgl_init_interface(cgl_t *h, double z0, double kprof)
{
    if (h && h->nbox) return fail(CGL_ERR_ARG, "a batch handle takes the cgl_box_* entry points");
    if (!h) return fail(CGL_ERR_ARG, "null handle");
    int rc = set_device(h);
    if (rc) return rc;
    cgl::CylIC ic{};
    ic.mode = 1;
    ic.z0 = z0;
    ic.kprof = kprof;
    h->cur = 0; // initial states are held in RB form in buffer 0
    h->state_layout = cgl::RB;
    h->primed = 0;
    h->stage = STAGE_NONE;
    h->phi_ok = 0;
    h->scal_stale = 0;
    (void)clear_range(h);
    dispatch(h, [&](auto impl) { decltype(impl)::init_cylinder(h, ic); });
    if ((rc = check_launch("init_interface"))) return rc;
    HIPCHK(hipStreamSynchronize(h->stream));
    return CGL_OK;
}

// Capture two steps from the current buffer parity into a graph (kept until
// the parity or the state layout changes).  Two steps return the buffers
// (populations, phi, particle masks) to the same parity, so one graph serves
// every replay.  The steps contain kernel launches only.
static int capture_two_steps(cgl_handle *h)
{
    const int key = h->cur | h->phicur << 1 | h->state_layout << 2 | h->ntcur << 3 | h->phi_ok << 4;
    if (h->gexec && h->gkey == key) return CGL_OK;
    if (h->gexec) {
        HIPCHK(hipGraphExecDestroy(h->gexec));
        h->gexec = nullptr;
        h->gkey = -1;
    }
    const int cur = h->cur, phicur = h->phicur, done = h->steps_done, ntcur = h->ntcur;
    const int phi_ok = h->phi_ok, stale = h->scal_stale;
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    enqueue_step(h, false);
    enqueue_step(h, false);
    hipError_t e = hipStreamEndCapture(h->stream, &g);
    h->cur = cur; // the captured steps have not run yet
    h->phicur = phicur;
    h->steps_done = done;
    h->ntcur = ntcur;
    h->phi_ok = phi_ok;
    h->scal_stale = stale;
    if (e != hipSuccess) return fail(CGL_ERR_HIP, "step capture failed: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&h->gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        h->gexec = nullptr;
        return fail(CGL_ERR_HIP, "graph instantiation failed: %s", hipGetErrorString(e));
    }
    h->gkey = key;
    return CGL_OK;
}

// Autotuner of the steps' launch shape: the x chunk (planes per workgroup),
// then the tile order over the XCDs, then the scatter-form step's tile
// height, each dimension with the others held at their current values.
// Results do not depend on any of them, so the first steps of a handle serve
// as the timing runs: each candidate advances TUNE_STEPS real steps between
// two events.
static const int TUNE_XC[] = {2, 4, 8, 16, 32, 64};
static const int TUNE_N = sizeof(TUNE_XC) / sizeof(TUNE_XC[0]);
static const int TUNE_REMAP[] = {0, 1, 2};
static const int TUNE_KY[] = {4, 8, 16};
static const int TUNE_STEPS = 2;
static const int TUNE_DIMS = 3;

static int tune_count(int dim) { return dim == 0 ? TUNE_N : 3; }
static int tune_value(int dim, int k) { return dim == 0 ? TUNE_XC[k] : (dim == 1 ? TUNE_REMAP[k] : TUNE_KY[k]); }

// whether candidate k of dimension dim applies to the handle
static bool tune_valid(const cgl_handle *h, int dim, int k)
{
    if (dim == 0) return k == 0 || TUNE_XC[k] < 2 * h->p.nx;
    if (dim == 1) return !h->remap_env && !h->nbox;
    return h->sc && !h->nbox && !h->ext && !h->tile_rows && cgl::sc_shape_ok(h->p.ny, h->p.nz, TUNE_KY[k]);
}

static void tune_set(cgl_handle *h, int dim, int v)
{
    if (dim == 0)
        h->xc = v;
    else if (dim == 1)
        (h->sc ? h->sc_remap : h->L.xcd_remap) = v;
    else
        h->ky = v;
}

// Move to the next candidate that applies; at the end of a dimension keep
// its best value and go on to the next; after the last one stop tuning.
static void tune_advance(cgl_handle *h)
{
    while (h->tuning) {
        if (h->tune_idx < tune_count(h->tune_dim)) {
            if (tune_valid(h, h->tune_dim, h->tune_idx)) return;
            ++h->tune_idx;
            continue;
        }
        if (h->tune_best_ms >= 0.f) tune_set(h, h->tune_dim, h->tune_best_val);
        h->tune_best_ms = -1.f;
        h->tune_idx = 0;
        if (++h->tune_dim == TUNE_DIMS) h->tuning = 0;
    }
}

static int tune_some(cgl_handle *h, int32_t *left)
{
    while (h->tuning && *left >= TUNE_STEPS) {
        // the layout conversion, the particles' start-up step and the scatter
        // form's first step (a moments pass) are not steady steps: run them
        // before a candidate's timing, so every candidate times the same work
        while (*left > 0 && needs_eager(h)) {
            enqueue_step(h);
            int rc = check_launch("step");
            if (rc) return rc;
            --*left;
        }
        if (*left < TUNE_STEPS) break;
        tune_advance(h);
        if (!h->tuning) break;
        const int v = tune_value(h->tune_dim, h->tune_idx);
        tune_set(h, h->tune_dim, v);
        HIPCHK(hipEventRecord(h->ev[0], h->stream));
        for (int t = 0; t < TUNE_STEPS; ++t) enqueue_step(h);
        HIPCHK(hipEventRecord(h->ev[1], h->stream));
        int rc = check_launch("step");
        if (rc) return rc;
        HIPCHK(hipEventSynchronize(h->ev[1]));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, h->ev[0], h->ev[1]));
        if (h->tune_best_ms < 0.f || ms < h->tune_best_ms) {
            h->tune_best_ms = ms;
            h->tune_best_val = v;
        }
        *left -= TUNE_STEPS;
        ++h->tune_idx;
        tune_advance(h);
    }
    return CGL_OK;
}

int cgl_step(cgl_t *h, int32_t nsteps)
{
    if (!h || nsteps < 0) return fail(CGL_ERR_ARG, "bad arguments");
    int rc = set_device(h);
    if (rc) return rc;
    if ((rc = no_open_stage(h))) return rc;
    h->stage = STAGE_NONE;
    if ((rc = tune_some(h, &nsteps))) return rc;
    if (h->use_graph && !h->tuning && nsteps >= 2) {
        // steady state: replay two captured steps (no host work per step).  The
        // layout conversion, the particles' start-up step and the scatter
        // form's first step run eagerly, and so does the scatter form's last
        // step, which writes the recolour scalars.
        while (nsteps > 0 && needs_eager(h)) {
            enqueue_step(h, nsteps == 1);
            if ((rc = check_launch("step"))) return rc;
            --nsteps;
        }
        const int pairs = (nsteps - (h->sc ? 1 : 0)) / 2;
        if (pairs > 0) {
            if ((rc = capture_two_steps(h))) return rc;
            for (int k = 0; k < pairs; ++k) HIPCHK(hipGraphLaunch(h->gexec, h->stream));
            h->steps_done += 2 * pairs;
            if (h->sc) {
                h->phi_ok = 1;
                h->scal_stale = 1;
            }
            nsteps -= 2 * pairs;
        }
    }
    for (int t = 0; t < nsteps; ++t) {
        enqueue_step(h, t == nsteps - 1);
        if ((rc = check_launch("step"))) return rc;
    }
    return CGL_OK;
}

int cgl_tuned_xchunk(cgl_t *h, int32_t *xchunk, int32_t *tuning)
{
    if (!h) return fail(CGL_ERR_ARG, "null handle");
    if (xchunk) *xchunk = h->xc;
    if (tuning) *tuning = h->tuning;
    return CGL_OK;
}

int cgl_tuned_params(cgl_t *h, int32_t *xchunk, int32_t *xcd_order, int32_t *tile_rows, int32_t *tuning)
{
    if (!h) return fail(CGL_ERR_ARG, "null handle");
    if (xchunk) *xchunk = h->xc;
    if (xcd_order) *xcd_order = h->sc ? h->sc_remap : h->L.xcd_remap;
    if (tile_rows) *tile_rows = h->sc ? h->ky : (h->p.precision == CGL_PREC_F64 ? cgl::KTY<double> : cgl::KTY<float>);
    if (tuning) *tuning = h->tuning;
    return CGL_OK;
}

int cgl_phase_field_form(cgl_t *h, int32_t *scatter)
{
    if (!h || !scatter) return fail(CGL_ERR_ARG, "null argument");
    *scatter = h->sc;
    return CGL_OK;
}

int cgl_calmacros(cgl_t *h, double *rho, double *rho_r, double *rho_b, double *ux, double *uy, double *uz, double *phi)
{
    if (h && h->nbox) return fail(CGL_ERR_ARG, "a batch handle takes the cgl_box_* entry points");
    if (!h) return fail(CGL_ERR_ARG, "null handle");
    int rc = set_device(h);
    if (rc) return rc;
    if ((rc = state_readable(h))) return rc;
    if ((rc = ensure_scratch(h, true, false, false))) return rc;
    dispatch(h, [&](auto impl) {
        if (h->ext)
            decltype(impl)::st_calmacros(h, true);
        else
            decltype(impl)::moments(h, true);
    });
    if ((rc = check_launch("moments"))) return rc;
    if (h->ext) h->stage = STAGE_CALMACROS; // the first stage of a staged step
    double *dst[7] = {rho, rho_r, rho_b, ux, uy, uz, phi};
    for (int k = 0; k < 7; ++k)
        if (dst[k]) HIPCHK(hipMemcpyAsync(dst[k], h->mac + k * h->N, h->N * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return CGL_OK;
}

int cgl_derivatives(cgl_t *h, double *gradx, double *grady, double *gradz, double *n_x, double *n_y, double *n_z,
                    int32_t *flag)
{
    if (h && h->nbox) return fail(CGL_ERR_ARG, "a batch handle takes the cgl_box_* entry points");
    if (!h) return fail(CGL_ERR_ARG, "null handle");
    int rc = set_device(h);
    if (rc) return rc;
    if ((rc = state_readable(h))) return rc;
    if ((rc = ensure_scratch(h, true, false, true))) return rc;
    h->stage = STAGE_NONE;
    dispatch(h, [&](auto impl) {
        if (h->ext) {
            decltype(impl)::derivatives_ext(h, h->mac, h->iflag);
        } else {
            decltype(impl)::moments(h, false);
            decltype(impl)::derivatives(h, h->mac, h->iflag);
        }
    });
    if ((rc = check_launch("derivatives"))) return rc;
    double *dst[6] = {gradx, grady, gradz, n_x, n_y, n_z};
    for (int k = 0; k < 6; ++k)
        if (dst[k]) HIPCHK(hipMemcpyAsync(dst[k], h->mac + k * h->N, h->N * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    if (flag) HIPCHK(hipMemcpyAsync(flag, h->iflag, h->N * sizeof(int), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return CGL_OK;
}

int cgl_sync(cgl_t *h)
{
    if (!h) return fail(CGL_ERR_ARG, "null handle");
    HIPCHK(hipStreamSynchronize(h->stream));
    return check_range(h);
}

int cgl_time_steps(cgl_t *h, int32_t nsteps, double *ms_total, double *ms_moments, double *ms_collide)
{
    if (!h || nsteps < 1) return fail(CGL_ERR_ARG, "bad arguments");
    int rc = set_device(h);
    if (rc) return rc;
    if ((rc = no_open_stage(h))) return rc;
    h->stage = STAGE_NONE;
    HIPCHK(hipStreamSynchronize(h->stream));
    // pass 1: the whole loop between two events (no per-kernel events inside)
    HIPCHK(hipEventRecord(h->ev[0], h->stream));
    for (int t = 0; t < nsteps; ++t) enqueue_step(h, t == nsteps - 1);
    HIPCHK(hipEventRecord(h->ev[1], h->stream));
    if ((rc = check_launch("step"))) return rc;
    HIPCHK(hipEventSynchronize(h->ev[1]));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, h->ev[0], h->ev[1]));
    if (ms_total) *ms_total = ms;
    // pass 2: per-kernel durations, one event pair around each launch
    double km = 0.0, kc = 0.0;
    for (int t = 0; t < nsteps && h->ext; ++t) {
        // with walls/particles: event pairs around the step's moments and
        // collide kernels, recorded by the step itself
        float a = 0.f, b = 0.f;
        h->tev = h->ev;
        enqueue_step(h);
        h->tev = nullptr;
        HIPCHK(hipStreamSynchronize(h->stream)); // (the scatter form records the moments pair last)
        HIPCHK(hipEventElapsedTime(&a, h->ev[0], h->ev[1]));
        HIPCHK(hipEventElapsedTime(&b, h->ev[2], h->ev[3]));
        km += a;
        kc += b;
    }
    // scatter form: the phase-field passes (a moments pass where needed, and
    // the finishing pass) count as moments, the step kernel as collide
    for (int t = 0; t < nsteps && !h->ext; ++t) {
        float a = 0.f, b = 0.f, c = 0.f;
        const bool scs = h->sc && h->state_layout == cgl::FS;
        HIPCHK(hipEventRecord(h->ev[0], h->stream));
        dispatch(h, [&](auto impl) {
            if (scs)
                decltype(impl)::sc_pre(h);
            else if (h->nbox)
                decltype(impl)::batch_moments(h);
            else
                decltype(impl)::moments(h, false);
        });
        HIPCHK(hipEventRecord(h->ev[1], h->stream));
        dispatch(h, [&](auto impl) {
            if (scs)
                decltype(impl)::sc_collide(h, t == nsteps - 1);
            else if (h->nbox)
                decltype(impl)::batch_collide(h);
            else
                decltype(impl)::collide(h);
        });
        HIPCHK(hipEventRecord(h->ev[2], h->stream));
        if (scs)
            dispatch(h, [&](auto impl) { decltype(impl)::sc_fin(h); });
        else
            h->phi_ok = 0;
        HIPCHK(hipEventRecord(h->ev[3], h->stream));
        h->cur ^= 1;
        h->phicur ^= 1;
        h->steps_done++;
        HIPCHK(hipEventSynchronize(h->ev[3]));
        HIPCHK(hipEventElapsedTime(&a, h->ev[0], h->ev[1]));
        HIPCHK(hipEventElapsedTime(&b, h->ev[1], h->ev[2]));
        HIPCHK(hipEventElapsedTime(&c, h->ev[2], h->ev[3]));
        km += a + (scs ? c : 0.f);
        kc += b;
    }
    if (ms_moments) *ms_moments = km;
    if (ms_collide) *ms_collide = kc;
    return check_range(h);
}

// Link-list storage for particles of radius up to rmax: the box k_lists
// scans, box_of(centre, R, 1), spans at most 2 ceil(R) + 5 nodes per axis,
// and each of its LSLICES slices at most a ceil(1 / LSLICES) share of it.
static int ensure_lists(cgl_handle *h, double rmax)
{
    const double w = 2.0 * ceil(rmax) + 5.0;
    if (!(w * w * w < 2147483647.0)) return fail(CGL_ERR_ARG, "particle radius %g too large", rmax);
    // the band box of the per-node kernels, box_of(centre, R, 3): at most
    // 2 ceil(R) + 9 nodes per axis, one thread per node
    const double wb = w + 4.0;
    h->band_blocks = (unsigned)((wb * wb * wb + cgl::PB - 1) / cgl::PB);
    // and the normals' box, box_of(centre, R, 4)
    const double wn = w + 6.0;
    h->nrm_blocks = (unsigned)((wn * wn * wn + cgl::PB - 1) / cgl::PB);
    const int cap = (int)((w * w * w + cgl::LSLICES - 1) / cgl::LSLICES);
    if (h->P.lout && cap <= h->P.lcap) return CGL_OK;
    unsigned **v[3] = {&h->P.lout, &h->P.lmask, &h->P.lin};
    for (unsigned **q : v) {
        if (*q) HIPCHK(hipFree(*q));
        *q = nullptr;
    }
    h->P.lcap = 0;
    const size_t bytes = (size_t)h->npar * cgl::LSLICES * (size_t)cap * sizeof(unsigned);
    for (unsigned **q : v) HIPCHK(dmalloc((void **)q, bytes, PZ_PART, h->stream));
    h->P.lcap = cap;
    return CGL_OK;
}

int cgl_set_particles(cgl_t *h, const double *center, const double *up, const double *rp, const double *rad,
                      const double *mass, const double *inertia, const double *cita, double fg)
{
    if (!h || !center || !rad || !mass || !inertia || !cita) return fail(CGL_ERR_ARG, "null argument");
    if (h->npar == 0) return fail(CGL_ERR_ARG, "the handle was created with npar = 0");
    int rc = set_device(h);
    if (rc) return rc;
    const int n3[3] = {h->p.nx, h->p.ny, h->p.nz};
    for (int k = 0; k < h->npar; ++k) {
        if (!(rad[k] > 0) || !(mass[k] > 0) || !(inertia[k] > 0)) return fail(CGL_ERR_ARG, "particle %d: bad radius or mass", k);
        for (int a = 0; a < 3; ++a)
            if (!h->E.wall[a][0] && !h->E.mirror[a][0] && 2.0 * (rad[k] + 4.0) + 3.0 > n3[a])
                return fail(CGL_ERR_ARG, "particle %d: box along periodic axis %d is smaller than the particle's reach", k, a);
    }
    const size_t np = (size_t)h->npar, b3 = 3 * np * sizeof(double), b1 = np * sizeof(double);
    // (npar, 3) component-major host arrays -> [npar][3] on the device
    std::vector<double> c3(3 * np), u3(3 * np, 0.0), w3(3 * np, 0.0);
    for (size_t p = 0; p < np; ++p)
        for (int k = 0; k < 3; ++k) {
            c3[3 * p + k] = center[k * np + p];
            if (up) u3[3 * p + k] = up[k * np + p];
            if (rp) w3[3 * p + k] = rp[k * np + p];
        }
    // the steps queued before this call read the particle arrays: they finish
    // first, and the copies run on the handle's stream (synchronised below,
    // before the host vectors go out of scope)
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipMemcpyAsync(h->P.center, c3.data(), b3, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->P.center_old, c3.data(), b3, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->P.center_prev, c3.data(), b3, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->P.center0, c3.data(), b3, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->P.up, u3.data(), b3, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->P.rp, w3.data(), b3, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->P.rad, rad, b1, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->P.mass, mass, b1, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->P.inertia, inertia, b1, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->P.cita, cita, b1, hipMemcpyHostToDevice, h->stream));
    h->P.fg = fg;
    // repulsion planes: wall faces sit half-way between ghost and first
    // layer; no Lagrange domain box, no symmetry rules (cgl_set_particle_geometry)
    for (int a = 0; a < 3; ++a) {
        h->P.rep_on[2 * a] = h->E.wall[a][0];
        h->P.rep_on[2 * a + 1] = h->E.wall[a][1];
        h->P.rep_plane[2 * a] = 1.5;
        h->P.rep_plane[2 * a + 1] = n3[a] - 0.5;
        h->P.dom_lo_on[a] = h->P.dom_lo_strict[a] = h->P.dom_hi_on[a] = 0;
        h->P.dom_lo[a] = h->P.dom_hi[a] = 0.0;
    }
    h->P.sym_rules = 0;
    if ((rc = drop_graph(h))) return rc;
    double rmax = 0.0;
    for (int k = 0; k < h->npar; ++k) rmax = rad[k] > rmax ? rad[k] : rmax;
    if ((rc = ensure_lists(h, rmax))) return rc;
    // set_walls of the start-up (S:212): fresh masks
    for (int b = 0; b < 2; ++b) HIPCHK(hipMemsetAsync(h->ntype[b], 0, h->N * sizeof(unsigned), h->stream));
    HIPCHK(hipMemsetAsync(h->walls_i, 0, h->N * sizeof(int), h->stream));
    HIPCHK(hipMemsetAsync(h->pin_i, 0, h->N * sizeof(int), h->stream));
    HIPCHK(hipMemsetAsync(h->pout_i, 0, h->N * sizeof(int), h->stream));
    h->ntcur = 0;
    dispatch(h, [&](auto impl) { decltype(impl)::build_masks(h, 0); });
    if ((rc = check_launch("set_walls"))) return rc;
    h->primed = 0;
    h->stage = STAGE_NONE;
    h->cold = 1;
    HIPCHK(hipStreamSynchronize(h->stream));
    return CGL_OK;
}

int cgl_set_particle_geometry(cgl_t *h, const double *rep_plane, const int32_t *rep_on, const double *dom_lo,
                              const double *dom_hi, const int32_t *dom_on, int32_t sym_rules)
{
    if (!h || !rep_plane || !rep_on || !dom_lo || !dom_hi || !dom_on) return fail(CGL_ERR_ARG, "null argument");
    if (h->npar == 0) return fail(CGL_ERR_ARG, "the handle was created with npar = 0");
    if (sym_rules != 0 && sym_rules != 1) return fail(CGL_ERR_ARG, "sym_rules must be 0 or 1");
    for (int f = 0; f < 6; ++f) {
        h->P.rep_plane[f] = rep_plane[f];
        h->P.rep_on[f] = rep_on[f] != 0;
    }
    for (int a = 0; a < 3; ++a) {
        h->P.dom_lo[a] = dom_lo[a];
        h->P.dom_hi[a] = dom_hi[a];
        h->P.dom_lo_on[a] = dom_on[3 * a] != 0;
        h->P.dom_lo_strict[a] = dom_on[3 * a + 1] != 0;
        h->P.dom_hi_on[a] = dom_on[3 * a + 2] != 0;
    }
    h->P.sym_rules = sym_rules;
    return drop_graph(h); // the captured steps carry the old geometry
}

int cgl_get_particles(cgl_t *h, double *center, double *up, double *rp, double *force, double *torque,
                      double *fcap, double *tcap, double *frep)
{
    if (!h) return fail(CGL_ERR_ARG, "null handle");
    if (h->npar == 0) return fail(CGL_ERR_ARG, "the handle was created with npar = 0");
    int rc = set_device(h);
    if (rc) return rc;
    const size_t np = (size_t)h->npar, b3 = 3 * np * sizeof(double);
    double *dst[8] = {center, up, rp, force, torque, fcap, tcap, frep};
    double *src[8] = {h->P.center, h->P.up, h->P.rp, h->P.fp, h->P.tp, h->P.ten, h->P.tent, h->P.frep};
    HIPCHK(hipStreamSynchronize(h->stream));
    std::vector<double> tmp(3 * np);
    for (int v = 0; v < 8; ++v) {
        if (!dst[v]) continue;
        HIPCHK(hipMemcpyAsync(tmp.data(), src[v], b3, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        for (size_t p = 0; p < np; ++p) // [npar][3] -> (npar, 3) component-major
            for (int k = 0; k < 3; ++k) dst[v][k * np + p] = tmp[3 * p + k];
    }
    return CGL_OK;
}

// ------------------------------------------------------------ staged steps
// One step with particles driven stage by stage, each stage the kernels
// cgl_step runs at that point, so that each can be checked against its
// reference routine.  cgl_calmacros opens a staged step.
extern "C++" {
static int stage_enter(cgl_handle *h, int prev, const char *name, bool particles)
{
    if (!h) return fail(CGL_ERR_ARG, "null handle");
    if (h->nbox || !h->ext)
        return fail(CGL_ERR_UNSUPPORTED, "%s: staged steps run on single boxes with walls or particles", name);
    if (particles && h->npar == 0) return fail(CGL_ERR_ARG, "%s: the handle was created with npar = 0", name);
    if (h->stage != prev)
        return fail(CGL_ERR_ARG,
                    "%s: stage out of order (a staged step is cgl_calmacros, cgl_capillary, cgl_momentum, "
                    "cgl_repulsive, cgl_move, cgl_set_walls, cgl_refill, cgl_collide_stream)",
                    name);
    return set_device(h);
}

template <class F>
static int run_stage(cgl_handle *h, int prev, int next, const char *name, F &&launch, bool particles = true)
{
    int rc = stage_enter(h, prev, name, particles);
    if (rc) return rc;
    dispatch(h, launch);
    if ((rc = check_launch(name))) return rc;
    HIPCHK(hipStreamSynchronize(h->stream));
    h->stage = next;
    return CGL_OK;
}
} // extern "C++"

int cgl_capillary(cgl_t *h)
{
    return run_stage(h, STAGE_CALMACROS, STAGE_CAPILLARY, "cgl_capillary", [&](auto impl) {
        decltype(impl)::st_lagrange(h);
        decltype(impl)::st_move(h, cgl::MV_CAP);
    });
}

int cgl_momentum(cgl_t *h)
{
    return run_stage(h, STAGE_CAPILLARY, STAGE_MOMENTUM, "cgl_momentum", [&](auto impl) {
        decltype(impl)::st_momentum(h);
        decltype(impl)::st_move(h, cgl::MV_MOM);
    });
}

int cgl_repulsive(cgl_t *h)
{
    return run_stage(h, STAGE_MOMENTUM, STAGE_REPULSIVE, "cgl_repulsive",
                     [&](auto impl) { decltype(impl)::st_move(h, cgl::MV_REP); });
}

int cgl_move(cgl_t *h)
{
    return run_stage(h, STAGE_REPULSIVE, STAGE_MOVE, "cgl_move",
                     [&](auto impl) { decltype(impl)::st_move(h, cgl::MV_MOVE); });
}

int cgl_set_walls(cgl_t *h)
{
    return run_stage(h, STAGE_MOVE, STAGE_SET_WALLS, "cgl_set_walls",
                     [&](auto impl) { decltype(impl)::st_set_walls(h); });
}

int cgl_refill(cgl_t *h)
{
    return run_stage(h, STAGE_SET_WALLS, STAGE_REFILL, "cgl_refill",
                     [&](auto impl) { decltype(impl)::st_refill(h); });
}

int cgl_collide_stream(cgl_t *h)
{
    const int prev = h && h->npar ? STAGE_REFILL : STAGE_CALMACROS;
    int rc = run_stage(
        h, prev, STAGE_NONE, "cgl_collide_stream", [&](auto impl) { decltype(impl)::st_collide_stream(h); }, false);
    if (rc) return rc;
    h->cur ^= 1;
    h->phicur ^= 1;
    h->steps_done++;
    return CGL_OK;
}

int cgl_particle_band(cgl_t *h, double *rho, double *rho_r, double *rho_b, double *ux, double *uy, double *uz)
{
    if (!h) return fail(CGL_ERR_ARG, "null handle");
    if (h->npar == 0) return fail(CGL_ERR_ARG, "the handle was created with npar = 0");
    int rc = set_device(h);
    if (rc) return rc;
    double *dst[6] = {rho, rho_r, rho_b, ux, uy, uz};
    for (int k = 0; k < 6; ++k)
        if (dst[k])
            HIPCHK(hipMemcpyAsync(dst[k], h->E.mac + k * h->N, h->N * sizeof(double), hipMemcpyDeviceToHost,
                                  h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return CGL_OK;
}

int cgl_get_particle_shb(cgl_t *h, double *shb)
{
    if (!h || !shb) return fail(CGL_ERR_ARG, "null argument");
    if (h->npar == 0) return fail(CGL_ERR_ARG, "the handle was created with npar = 0");
    int rc = set_device(h);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipMemcpyAsync(shb, h->P.shb, (size_t)h->npar * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return CGL_OK;
}

__global__ void k_masks_out(const unsigned *nt, size_t N, int *walls_all)
{
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n < N) walls_all[n] = (nt[n] & cgl::NT_SOLID) ? 10 : 0;
}

int cgl_download_masks(cgl_t *h, int32_t *walls, int32_t *walls_all, int32_t *pin, int32_t *pout)
{
    if (!h) return fail(CGL_ERR_ARG, "null handle");
    if (h->npar == 0) return fail(CGL_ERR_ARG, "the handle was created with npar = 0");
    int rc = set_device(h);
    if (rc) return rc;
    const size_t bytes = h->N * sizeof(int);
    if (walls) HIPCHK(hipMemcpyAsync(walls, h->walls_i, bytes, hipMemcpyDeviceToHost, h->stream));
    if (pin) HIPCHK(hipMemcpyAsync(pin, h->pin_i, bytes, hipMemcpyDeviceToHost, h->stream));
    if (pout) HIPCHK(hipMemcpyAsync(pout, h->pout_i, bytes, hipMemcpyDeviceToHost, h->stream));
    if (walls_all) {
        if ((rc = ensure_scratch(h, false, false, true))) return rc;
        // between set_walls and the end of a staged step the new masks are the spare buffer
        const int nt = h->stage >= STAGE_SET_WALLS ? h->ntcur ^ 1 : h->ntcur;
        k_masks_out<<<(unsigned)((h->N + 255) / 256), 256, 0, h->stream>>>(h->ntype[nt], h->N, h->iflag);
        if ((rc = check_launch("masks"))) return rc;
        HIPCHK(hipMemcpyAsync(walls_all, h->iflag, bytes, hipMemcpyDeviceToHost, h->stream));
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    return CGL_OK;
}

int cgl_diagnostics(cgl_t *h, double *sums, double *slice_red)
{
    if (!h || !sums) return fail(CGL_ERR_ARG, "null argument");
    if (h->nbox) return fail(CGL_ERR_ARG, "a batch handle takes the cgl_box_* entry points");
    int rc = set_device(h);
    if (rc) return rc;
    if ((rc = state_readable(h))) return rc;
    if ((rc = ensure_scratch(h, true, false, false))) return rc;
    dispatch(h, [&](auto impl) {
        if (h->ext)
            decltype(impl)::moments_ext(h, true);
        else
            decltype(impl)::moments(h, true);
    });
    if ((rc = check_launch("moments"))) return rc;
    const unsigned per_block = 16384;
    const unsigned nblocks = (unsigned)((h->N + per_block - 1) / per_block);
    double *part = nullptr;
    HIPCHK(dmalloc((void **)&part, ((size_t)nblocks + 1) * cgl::NDIAG * sizeof(double) + h->p.nz * sizeof(double), PZ_SCR, h->stream));
    double *out = part + (size_t)nblocks * cgl::NDIAG, *area = out + cgl::NDIAG;
    const cgl::Ext E = h->npar ? [&] { cgl::Ext e = h->E; e.ntype = h->ntype[h->ntcur]; return e; }() : h->E;
    cgl::k_reduce<<<nblocks, cgl::RED_B, 0, h->stream>>>(h->L, E, h->mac, part, per_block);
    cgl::k_reduce_final<<<1, 64, 0, h->stream>>>(part, nblocks, out);
    if (slice_red) cgl::k_slice_red<<<h->p.nz, cgl::RED_B, 0, h->stream>>>(h->L, E, h->mac, area);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(sums, out, cgl::NDIAG * sizeof(double), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess && slice_red)
        e = hipMemcpyAsync(slice_red, area, h->p.nz * sizeof(double), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(part);
    if (e != hipSuccess) return fail(CGL_ERR_HIP, "diagnostics failed: %s", hipGetErrorString(e));
    return CGL_OK;
}

// ------------------------------------------- diagnostics without allocation
// The scratch of cgl_diagnostics_noalloc: a file-scope device array of
// CGL_NOALLOC_SCRATCH_DOUBLES doubles (4 MiB), one copy per GPU, held until
// the process ends.  Host code never passes the array's name to a kernel or a
// copy (on the host it is only a stand-in for the symbol): it takes the
// device address hipGetSymbolAddress reports, looked up once per device with
// that device current and kept per device.
static __device__ __align__(256) double scratch[CGL_NOALLOC_SCRATCH_DOUBLES];

namespace {
constexpr int SCR_MAXDEV = 64;
std::mutex g_scr_mu;
double *g_scr[SCR_MAXDEV] = {};

// the device address of `scratch` on the current device `dev`
int scratch_ptr(int dev, double **p)
{
    if (dev < 0 || dev >= SCR_MAXDEV) return fail(CGL_ERR_UNSUPPORTED, "device %d beyond the scratch table", dev);
    std::lock_guard<std::mutex> lk(g_scr_mu);
    if (!g_scr[dev]) {
        void *a = nullptr;
        const hipError_t e = hipGetSymbolAddress(&a, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !a)
            return fail(CGL_ERR_HIP, "scratch address on device %d: %s", dev, hipGetErrorString(e));
        g_scr[dev] = (double *)a;
    }
    *p = g_scr[dev];
    return CGL_OK;
}
} // namespace

int cgl_diagnostics_noalloc(cgl_t *h, double *sums, double *slice_red)
{
    if (!h || !sums) return fail(CGL_ERR_ARG, "null argument");
    if (h->nbox) return fail(CGL_ERR_ARG, "a batch handle takes the cgl_box_* entry points");
    int rc = set_device(h);
    if (rc) return rc;
    if ((rc = state_readable(h))) return rc;
    size_t need = 0;
    dispatch(h, [&](auto impl) { need = decltype(impl)::diag_scratch_doubles(h); });
    if (need > (size_t)CGL_NOALLOC_SCRATCH_DOUBLES)
        return fail(CGL_ERR_UNSUPPORTED,
                    "box %dx%dx%d needs %zu doubles of diagnostics scratch; the static buffer holds %d "
                    "(use cgl_diagnostics)",
                    h->p.nx, h->p.ny, h->p.nz, need, CGL_NOALLOC_SCRATCH_DOUBLES);
    double *scr = nullptr;
    if ((rc = scratch_ptr(h->p.device, &scr))) return rc;
    dispatch(h, [&](auto impl) { decltype(impl)::diag_fused(h, scr); });
    if ((rc = check_launch("diagnostics"))) return rc;
    HIPCHK(hipMemcpyAsync(sums, scr, cgl::NDIAG * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    if (slice_red)
        HIPCHK(hipMemcpyAsync(slice_red, scr + 8, h->p.nz * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return CGL_OK;
}

// ------------------------------------------------------------------ batches

int cgl_batch_create(cgl_t **out, const cgl_params *p, int32_t nbox, const cgl_box *boxes)
{
    if (!out || !p || !boxes || nbox < 1) return fail(CGL_ERR_ARG, "bad arguments");
    *out = nullptr;
    if (p->precision < CGL_PREC_F64 || p->precision > CGL_PREC_F16)
        return fail(CGL_ERR_ARG, "unknown precision %d", p->precision);
    if (p->layout != CGL_LAYOUT_RB && p->layout != CGL_LAYOUT_FS) return fail(CGL_ERR_ARG, "unknown layout %d", p->layout);
    for (int f = 0; f < 6; ++f)
        if (p->bc[f] != CGL_BC_PERIODIC) return fail(CGL_ERR_UNSUPPORTED, "batched boxes are periodic");
    if (p->npar != 0) return fail(CGL_ERR_UNSUPPORTED, "batched boxes carry no particles");
    if (!(p->rho0 > 0)) return fail(CGL_ERR_ARG, "rho0 must be > 0");
    if (p->tile_rows != 0 && p->tile_rows != 8) return fail(CGL_ERR_ARG, "batches run tiles of 8 rows (tile_rows 0 or 8)");
    const double eb = addressed_bytes(p);
    size_t total = 0, nmax = 0;
    for (int b = 0; b < nbox; ++b) {
        const cgl_box &B = boxes[b];
        if (B.nx < 1 || B.ny < 1 || B.nz < 1) return fail(CGL_ERR_ARG, "box %d: must be at least 1x1x1", b);
        if (!(B.nu_r > 0) || !(B.nu_b > 0)) return fail(CGL_ERR_ARG, "box %d: nu_r, nu_b must be > 0", b);
        if (B.reserved != 0) return fail(CGL_ERR_ARG, "box %d: reserved field must be 0", b);
        const size_t n = (size_t)B.nx * B.ny * B.nz;
        if ((double)n * eb > 4294967296.0) return fail(CGL_ERR_UNSUPPORTED, "box %d exceeds the 4 GiB per-array addressing", b);
        total += n;
        if (n > nmax) nmax = n;
    }
    cgl_params q = *p;
    q.nx = boxes[0].nx;
    q.ny = boxes[0].ny;
    q.nz = boxes[0].nz;
    q.nu_r = boxes[0].nu_r;
    q.nu_b = boxes[0].nu_b;
    int rc = alloc_handle(out, &q, total, 0);
    if (rc) return rc;
    cgl_handle *h = *out;
    h->tuning = 0;
    h->xc = p->xchunk > 0 ? p->xchunk : 8;
    h->scratch_n = nmax;
    h->hboxes = new (std::nothrow) cgl::BoxDev[nbox];
    if (!h->hboxes) {
        cgl_destroy(h);
        *out = nullptr;
        return fail(CGL_ERR_ARG, "out of host memory");
    }
    h->nbox = nbox;
    unsigned long long base = 0, t1 = 0, t2 = 0;
    for (int b = 0; b < nbox; ++b) {
        cgl::BoxDev &D = h->hboxes[b];
        memset(&D, 0, sizeof D);
        D.L = h->L;
        D.L.nx = boxes[b].nx;
        D.L.ny = boxes[b].ny;
        D.L.nz = boxes[b].nz;
        D.L.N = (long long)boxes[b].nx * boxes[b].ny * boxes[b].nz;
        D.L.nu_r = boxes[b].nu_r;
        D.L.nu_b = boxes[b].nu_b;
        D.base = base;
        D.gz = (boxes[b].nz + cgl::TZ - 1) / cgl::TZ;
        D.gy = (boxes[b].ny + cgl::TY - 1) / cgl::TY;
        D.gx2 = (boxes[b].nx + h->xc - 1) / h->xc;
        const int kty = p->precision == CGL_PREC_F64 ? cgl::KTY<double> : cgl::KTY<float>;
        D.gy2 = (boxes[b].ny + kty - 1) / kty;
        D.gx1 = (boxes[b].nx + cgl::K1_XC - 1) / cgl::K1_XC;
        D.t1 = (unsigned)t1;
        D.t2 = (unsigned)t2;
        t1 += (unsigned long long)D.gz * D.gy * D.gx1;
        t2 += (unsigned long long)D.gz * D.gy2 * D.gx2;
        base += (unsigned long long)D.L.N;
    }
    if (t1 >= (1ull << 31) || t2 >= (1ull << 31)) {
        cgl_destroy(h);
        *out = nullptr;
        return fail(CGL_ERR_UNSUPPORTED, "batch exceeds the launch's workgroup count");
    }
    h->tiles1 = (unsigned)t1;
    h->tiles2 = (unsigned)t2;
    hipError_t e = dmalloc((void **)&h->dboxes, nbox * sizeof(cgl::BoxDev), PZ_TAB, h->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(h->dboxes, h->hboxes, nbox * sizeof(cgl::BoxDev), hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) {
        rc = fail(CGL_ERR_HIP, "box table: %s", hipGetErrorString(e));
        cgl_destroy(h);
        *out = nullptr;
        return rc;
    }
    if ((rc = setup_scatter(h))) {
        cgl_destroy(h);
        *out = nullptr;
    }
    return rc;
}

// A single-box handle that aliases box `b` of a batch: same buffers, offset
// to the box's nodes.  Scratch is the batch's (sized for its largest box).
static int box_view(cgl_handle *h, int32_t b, cgl_handle *v)
{
    if (!h || !h->nbox) return fail(CGL_ERR_ARG, "not a batch handle");
    if (b < 0 || b >= h->nbox) return fail(CGL_ERR_ARG, "box %d out of range", b);
    int rc = set_device(h);
    if (rc) return rc;
    if ((rc = ensure_scratch(h, true, true, true))) return rc;
    *v = *h;
    const cgl::BoxDev &D = h->hboxes[b];
    v->nbox = 0;
    v->hboxes = nullptr;
    v->dboxes = nullptr;
    v->gexec = nullptr;
    v->gkey = -1;
    v->L = D.L;
    v->N = (size_t)D.L.N;
    v->p.nx = D.L.nx;
    v->p.ny = D.L.ny;
    v->p.nz = D.L.nz;
    v->p.nu_r = D.L.nu_r;
    v->p.nu_b = D.L.nu_b;
    const size_t base = (size_t)D.base;
    const size_t ns = h->p.precision == CGL_PREC_F64 ? 6 * sizeof(double) : 4 * sizeof(float);
    for (int k = 0; k < 2; ++k) {
        for (int c = 0; c < 2; ++c)
            if (h->pops[k][c]) v->pops[k][c] = (char *)h->pops[k][c] + (size_t)cgl::Q * base * h->esize;
        if (h->fsf[k]) v->fsf[k] = (char *)h->fsf[k] + (size_t)cgl::Q * base * h->esize;
        if (h->fss[k]) v->fss[k] = (char *)h->fss[k] + base * ns;
        v->phi[k] = (char *)h->phi[k] + base * h->phisize;
    }
    return CGL_OK;
}

static void box_back(cgl_handle *h, const cgl_handle *v)
{
    h->cur = v->cur;
    h->phicur = v->phicur;
    h->state_layout = v->state_layout;
    h->phi_ok = v->phi_ok; // an upload into one box invalidates the batch's carried phi
}

int cgl_box_upload_state(cgl_t *h, int32_t box, const double *ri, const double *bi)
{
    cgl_handle v;
    int rc = box_view(h, box, &v);
    if (rc) return rc;
    if (h->steps_done > 0) return fail(CGL_ERR_ARG, "upload every box of a batch before its first step");
    if ((rc = cgl_upload_state(&v, ri, bi))) return rc;
    box_back(h, &v);
    return CGL_OK;
}

int cgl_box_init_cylinder(cgl_t *h, int32_t box, double xc, double yc, double r0, double kprof, int32_t nmodes,
                          const double *amp)
{
    cgl_handle v;
    int rc = box_view(h, box, &v);
    if (rc) return rc;
    if (h->steps_done > 0) return fail(CGL_ERR_ARG, "initialise every box of a batch before its first step");
    if ((rc = cgl_init_cylinder(&v, xc, yc, r0, kprof, nmodes, amp))) return rc;
    box_back(h, &v);
    return CGL_OK;
}

int cgl_box_download_state(cgl_t *h, int32_t box, double *ri, double *bi)
{
    cgl_handle v;
    int rc = box_view(h, box, &v);
    if (rc) return rc;
    return cgl_download_state(&v, ri, bi);
}

int cgl_box_calmacros(cgl_t *h, int32_t box, double *rho, double *rho_r, double *rho_b, double *ux, double *uy,
                      double *uz, double *phi)
{
    cgl_handle v;
    int rc = box_view(h, box, &v);
    if (rc) return rc;
    return cgl_calmacros(&v, rho, rho_r, rho_b, ux, uy, uz, phi);
}

// ------------------------------------------------------------ checkpoints
// A checkpoint holds the device state as the handle keeps it (storage
// precision and layout, shifted values, both phi buffers, masks, the
// particles' band fields and state) plus the step bookkeeping, so a handle
// of the same configuration resumes bit for bit.  Little-endian host order,
// a fixed header, then the sections in a fixed order.

struct CkptHeader {
    char magic[8];
    int32_t version;
    int32_t nx, ny, nz, precision, layout, npar, nbox;
    int32_t bc[6];
    int32_t cur, phicur, state_layout, ntcur, phi_ok, primed, cold, xc, tuning, tune_idx, tune_best_val;
    int32_t tune_dim, ky, sc_remap, xcd_remap;
    float tune_best_ms;
    int64_t N;
    int64_t steps_done;
    double sigma, rho0, nu_r, nu_b, beta;
};
static const char CKPT_MAGIC[8] = {'C', 'G', 'L', 'C', 'K', 'P', 'T', '1'};
static const int CKPT_VERSION = 3;

// a batch's boxes, written after the header: a load must find the same
// boxes (sizes and viscosities) in the same order
struct CkptBox {
    int32_t nx, ny, nz, pad;
    double nu_r, nu_b;
};

// the particles' host-side scalars, written as one block
struct CkptPart {
    double fg, rep_plane[6], dom_lo[3], dom_hi[3];
    int32_t rep_on[6], dom_lo_on[3], dom_lo_strict[3], dom_hi_on[3], sym_rules;
};

// device buffers of a checkpoint, in file order
static std::vector<std::pair<void *, size_t>> ckpt_sections(cgl_handle *h)
{
    std::vector<std::pair<void *, size_t>> v;
    const size_t N = h->N, pops = (size_t)cgl::Q * N * h->esize;
    if (h->state_layout == cgl::RB) {
        v.push_back({h->pops[h->cur][0], pops});
        v.push_back({h->pops[h->cur][1], pops});
    } else {
        const size_t ns = h->p.precision == CGL_PREC_F64 ? 6 * sizeof(double) : 4 * sizeof(float);
        v.push_back({h->fsf[h->cur], pops});
        v.push_back({h->fss[h->cur], N * ns});
        if (h->ext) { // the side arrays of the solid nodes
            if (h->p.precision == CGL_PREC_F64) {
                v.push_back({h->pops[h->cur][0], pops});
                v.push_back({h->pops[h->cur][1], pops});
            } else {
                v.push_back({h->pops[h->cur][0], (size_t)cgl::Q * N * sizeof(float)});
            }
        }
    }
    v.push_back({h->phi[0], N * h->phisize});
    v.push_back({h->phi[1], N * h->phisize});
    if (h->npar) {
        v.push_back({h->ntype[0], N * sizeof(unsigned)});
        v.push_back({h->ntype[1], N * sizeof(unsigned)});
        v.push_back({h->walls_i, N * sizeof(int)});
        v.push_back({h->pin_i, N * sizeof(int)});
        v.push_back({h->pout_i, N * sizeof(int)});
        v.push_back({h->E.mac, 6 * N * sizeof(double)});
        v.push_back({h->E.der, 6 * N * sizeof(double)});
        const size_t n3 = 3 * (size_t)h->npar * sizeof(double), n1 = (size_t)h->npar * sizeof(double);
        double *vec[11] = {h->P.center, h->P.center_old, h->P.center_prev, h->P.center0, h->P.up, h->P.rp,
                           h->P.fp, h->P.tp, h->P.frep, h->P.ten, h->P.tent};
        for (double *q : vec) v.push_back({q, n3});
        double *sc[5] = {h->P.rad, h->P.mass, h->P.inertia, h->P.cita, h->P.shb};
        for (double *q : sc) v.push_back({q, n1});
    }
    return v;
}

static const size_t CKPT_CHUNK = 64u << 20;

int cgl_checkpoint_save(cgl_t *h, const char *path)
{
    if (!h || !path) return fail(CGL_ERR_ARG, "null argument");
    int rc = set_device(h);
    if (rc) return rc;
    if ((rc = state_readable(h))) return rc;
    HIPCHK(hipStreamSynchronize(h->stream));
    CkptHeader H;
    memset(&H, 0, sizeof H);
    memcpy(H.magic, CKPT_MAGIC, 8);
    H.version = CKPT_VERSION;
    H.nx = h->p.nx; H.ny = h->p.ny; H.nz = h->p.nz;
    H.precision = h->p.precision; H.layout = h->p.layout; H.npar = h->npar; H.nbox = h->nbox;
    for (int f = 0; f < 6; ++f) H.bc[f] = h->p.bc[f];
    H.cur = h->cur; H.phicur = h->phicur; H.state_layout = h->state_layout; H.ntcur = h->ntcur;
    H.phi_ok = h->phi_ok; H.primed = h->primed; H.cold = h->cold; H.xc = h->xc; H.tuning = h->tuning;
    H.tune_idx = h->tune_idx; H.tune_best_val = h->tune_best_val; H.tune_best_ms = h->tune_best_ms;
    H.tune_dim = h->tune_dim; H.ky = h->ky; H.sc_remap = h->sc_remap; H.xcd_remap = h->L.xcd_remap;
    H.N = (int64_t)h->N; H.steps_done = h->steps_done;
    H.sigma = h->p.sigma; H.rho0 = h->p.rho0; H.nu_r = h->p.nu_r; H.nu_b = h->p.nu_b; H.beta = h->p.beta;
    FILE *f = fopen(path, "wb");
    if (!f) return fail(CGL_ERR_ARG, "cannot open %s for writing", path);
    bool ok = fwrite(&H, sizeof H, 1, f) == 1;
    for (int b = 0; ok && b < h->nbox; ++b) {
        const cgl::Lat &Lb = h->hboxes[b].L;
        const CkptBox cb{Lb.nx, Lb.ny, Lb.nz, 0, Lb.nu_r, Lb.nu_b};
        ok = fwrite(&cb, sizeof cb, 1, f) == 1;
    }
    if (ok && h->npar) {
        CkptPart Pp;
        memset(&Pp, 0, sizeof Pp);
        Pp.fg = h->P.fg;
        for (int k = 0; k < 6; ++k) { Pp.rep_plane[k] = h->P.rep_plane[k]; Pp.rep_on[k] = h->P.rep_on[k]; }
        for (int a = 0; a < 3; ++a) {
            Pp.dom_lo[a] = h->P.dom_lo[a]; Pp.dom_hi[a] = h->P.dom_hi[a];
            Pp.dom_lo_on[a] = h->P.dom_lo_on[a]; Pp.dom_lo_strict[a] = h->P.dom_lo_strict[a];
            Pp.dom_hi_on[a] = h->P.dom_hi_on[a];
        }
        Pp.sym_rules = h->P.sym_rules;
        ok = fwrite(&Pp, sizeof Pp, 1, f) == 1;
    }
    std::vector<char> buf;
    hipError_t e = hipSuccess;
    for (auto &sec : ckpt_sections(h)) {
        for (size_t off = 0; ok && e == hipSuccess && off < sec.second; off += CKPT_CHUNK) {
            const size_t n = sec.second - off < CKPT_CHUNK ? sec.second - off : CKPT_CHUNK;
            buf.resize(n);
            e = hipMemcpyAsync(buf.data(), (char *)sec.first + off, n, hipMemcpyDeviceToHost, h->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
            if (e == hipSuccess) ok = fwrite(buf.data(), 1, n, f) == n;
        }
    }
    ok = (fclose(f) == 0) && ok;
    if (e != hipSuccess) return fail(CGL_ERR_HIP, "checkpoint copy failed: %s", hipGetErrorString(e));
    if (!ok) return fail(CGL_ERR_ARG, "writing %s failed", path);
    return CGL_OK;
}

int cgl_checkpoint_load(cgl_t *h, const char *path)
{
    if (!h || !path) return fail(CGL_ERR_ARG, "null argument");
    int rc = set_device(h);
    if (rc) return rc;
    FILE *f = fopen(path, "rb");
    if (!f) return fail(CGL_ERR_ARG, "cannot open %s", path);
    CkptHeader H;
    if (fread(&H, sizeof H, 1, f) != 1 || memcmp(H.magic, CKPT_MAGIC, 8) != 0 || H.version != CKPT_VERSION) {
        fclose(f);
        return fail(CGL_ERR_ARG, "%s is not a libcgl checkpoint of this version", path);
    }
    bool same = H.nx == h->p.nx && H.ny == h->p.ny && H.nz == h->p.nz && H.precision == h->p.precision &&
                H.layout == h->p.layout && H.npar == h->npar && H.nbox == h->nbox && H.N == (int64_t)h->N &&
                H.sigma == h->p.sigma && H.rho0 == h->p.rho0 && H.beta == h->p.beta &&
                (h->nbox || (H.nu_r == h->p.nu_r && H.nu_b == h->p.nu_b));
    for (int k = 0; k < 6; ++k) same = same && H.bc[k] == h->p.bc[k];
    const bool layout_ok = H.state_layout == cgl::RB ? (h->p.layout == CGL_LAYOUT_RB || H.cur == 0)
                                                     : h->p.layout == CGL_LAYOUT_FS;
    // a batch's tile tables were built for its x chunk at creation; a single
    // box takes any chunk, and the autotuner's state must index its tables;
    // the tile height must suit the box
    const bool tune_ok =
        (H.tuning == 0 || (H.tuning == 1 && H.tune_dim >= 0 && H.tune_dim < TUNE_DIMS && H.tune_idx >= 0 &&
                           H.tune_idx <= tune_count(H.tune_dim))) &&
        (H.ky == 4 || H.ky == 8 || H.ky == 16) && (!h->sc || cgl::sc_shape_ok(h->p.ny, h->p.nz, H.ky)) &&
        H.sc_remap >= 0 && H.sc_remap <= 2 && H.xcd_remap >= 0 && H.xcd_remap <= 2;
    if (!same || !layout_ok || H.cur < 0 || H.cur > 1 || H.phicur < 0 || H.phicur > 1 || H.ntcur < 0 ||
        H.ntcur > 1 || H.xc < 1 || H.xc > h->p.nx + 64 || (h->nbox && (H.xc != h->xc || H.tuning)) || !tune_ok) {
        fclose(f);
        return fail(CGL_ERR_ARG, "%s holds a different configuration", path);
    }
    for (int b = 0; b < h->nbox; ++b) {
        CkptBox cb;
        const cgl::Lat &Lb = h->hboxes[b].L;
        if (fread(&cb, sizeof cb, 1, f) != 1 || cb.nx != Lb.nx || cb.ny != Lb.ny || cb.nz != Lb.nz ||
            cb.nu_r != Lb.nu_r || cb.nu_b != Lb.nu_b) {
            fclose(f);
            return fail(CGL_ERR_ARG, "%s holds a batch of other boxes (box %d)", path, b);
        }
    }
    CkptPart Pp;
    if (h->npar && fread(&Pp, sizeof Pp, 1, f) != 1) {
        fclose(f);
        return fail(CGL_ERR_ARG, "%s is truncated", path);
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    h->cur = H.cur;
    h->state_layout = H.state_layout;
    std::vector<char> buf;
    bool ok = true;
    hipError_t e = hipSuccess;
    for (auto &sec : ckpt_sections(h)) {
        for (size_t off = 0; ok && e == hipSuccess && off < sec.second; off += CKPT_CHUNK) {
            const size_t n = sec.second - off < CKPT_CHUNK ? sec.second - off : CKPT_CHUNK;
            buf.resize(n);
            ok = fread(buf.data(), 1, n, f) == n;
            if (ok) e = hipMemcpyAsync((char *)sec.first + off, buf.data(), n, hipMemcpyHostToDevice, h->stream);
            if (ok && e == hipSuccess) e = hipStreamSynchronize(h->stream);
        }
    }
    fclose(f);
    if (e != hipSuccess) return fail(CGL_ERR_HIP, "checkpoint copy failed: %s", hipGetErrorString(e));
    if (!ok) {
        h->phi_ok = 0; // the state is incomplete: take phi from a moments pass
        return fail(CGL_ERR_ARG, "%s is truncated", path);
    }
    h->phicur = H.phicur; h->ntcur = H.ntcur; h->phi_ok = H.phi_ok; h->primed = H.primed; h->cold = H.cold;
    h->xc = H.xc; h->tuning = H.tuning; h->tune_idx = H.tune_idx; h->tune_best_val = H.tune_best_val;
    h->tune_best_ms = H.tune_best_ms; h->tune_dim = H.tune_dim; h->steps_done = (int)H.steps_done; h->scal_stale = 0;
    if (!h->tile_rows) h->ky = H.ky;
    if (!h->remap_env) {
        h->sc_remap = H.sc_remap;
        h->L.xcd_remap = H.xcd_remap;
    }
    (void)clear_range(h);
    h->stage = STAGE_NONE;
    if (h->npar) {
        h->P.fg = Pp.fg;
        for (int k = 0; k < 6; ++k) { h->P.rep_plane[k] = Pp.rep_plane[k]; h->P.rep_on[k] = Pp.rep_on[k]; }
        for (int a = 0; a < 3; ++a) {
            h->P.dom_lo[a] = Pp.dom_lo[a]; h->P.dom_hi[a] = Pp.dom_hi[a];
            h->P.dom_lo_on[a] = Pp.dom_lo_on[a]; h->P.dom_lo_strict[a] = Pp.dom_lo_strict[a];
            h->P.dom_hi_on[a] = Pp.dom_hi_on[a];
        }
        h->P.sym_rules = Pp.sym_rules;
        // the link lists are derived from the restored masks and centres
        std::vector<double> rad((size_t)h->npar);
        HIPCHK(hipMemcpyAsync(rad.data(), h->P.rad, rad.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        double rmax = 0.0;
        for (double r : rad) rmax = r > rmax ? r : rmax;
        if ((rc = drop_graph(h))) return rc;
        if ((rc = ensure_lists(h, rmax))) return rc;
        dispatch(h, [&](auto impl) { decltype(impl)::build_lists(h, h->ntcur); });
        if ((rc = check_launch("link lists"))) return rc;
        HIPCHK(hipStreamSynchronize(h->stream));
    }
    return drop_graph(h);
}

int cgl_device_name(cgl_t *h, char *buf, int32_t len)
{
    if (!h || !buf || len < 1) return fail(CGL_ERR_ARG, "bad arguments");
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, h->p.device));
    snprintf(buf, (size_t)len, "%s (%s)", prop.name, prop.gcnArchName);
    return CGL_OK;
}

} // extern "C"
