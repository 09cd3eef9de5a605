// cgl_kernels.hpp -- HIP kernels of the per-timestep lattice update for
// gfx950, templated on storage precision.
//
// Device data (all structure-of-arrays, node index n = (x*ny + y)*nz + z,
// z fastest as in the reference's (zl, yl, xl) arrays):
//   pops   r[i*N + n], b[i*N + n], i = 0..18, two buffers (A/B ping-pong).
//          They hold the post-collision, post-recolour populations of the
//          last step, i.e. the state right before the reference's streaming
//          (Source1.f90 S:249-250); a pull gather from x - c_i is streaming.
//   phi    phase field, two buffers: the step's phi, and the phi that the
//          stored populations were shifted with (FP32/FP16 storage only).
//
// Per step (S:216-251 without particles):
//   K1 k_moments(_fsx) pull-gather -> rho_R, rho_B -> phi     (calmacros, S:770-831)
//   K2 k_collide  x-marching tile, LDS rings of phi and n    (derivatives S:452-597,
//                 -> grad phi, n, curvature, CSF force,        collision S:599-695,
//                 MRT + Guo, recolour, write pops              Recolor S:697-732,
//                                                              streaming S:735-768)
// The factorised layout (FS, below) keeps the recolour scalars of planes
// x-1..x+1 in an LDS ring in both kernels.
//
// FP64 ("plain") keeps the oracle's exact operation order and is compiled
// without FMA contraction, so it reproduces the FP64 CPU path bit for bit.
// FP32/FP16 storage keeps each population shifted by w_i*rho0*a of the node
// that will read it, a = (1 + phi)/2 for red and (1 - phi)/2 for blue, with
// phi of the step that wrote it: the writer has that phi in its LDS ring and
// the reader needs only its own node's value.  Collision works on the
// deviations f - w_i*rho0 in FP32 with compensated moment sums.
#pragma once

#include <type_traits>

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cgl_lattice.hpp"

namespace cgl {

constexpr double EVAP_LIM = 1.0e-8; // S:16
constexpr int TZ = 64;              // workgroup extent along z (one wavefront)
constexpr int TY = 4;               // workgroup extent along y
// Workgroup extent along y of the fused collide kernel: 8 rows for FP32/FP16
// arithmetic (two 512-thread workgroups fit a CU's LDS), 4 for FP64, whose
// larger LDS rings would leave room for only one 512-thread workgroup.
template <class R> constexpr int KTY = sizeof(R) == 8 ? 4 : 8;
// Waves per SIMD the fused collide kernel is compiled for: 4 for FP32/FP16,
// so that it stays within 128 VGPRs and two 512-thread workgroups share a CU
// (the wall/particle variant spills about 20 registers for it and still runs
// 1.2x faster than at one workgroup per CU); no bound for FP64, which runs
// 256-thread workgroups.
template <class R> constexpr int KWAVES = sizeof(R) == 8 ? 1 : 4;

struct Lat {
    int nx, ny, nz;
    int xcd_remap; // 1: hand each XCD a contiguous range of tiles
    long long N;
    double rho0, sigma, beta, nu_r, nu_b;
};

// Element access through a 32-bit byte offset from a wave-uniform base: every
// per-direction array is below 4 GiB (checked at creation), so loads and
// stores use the scalar-base + 32-bit vector-offset addressing of gfx950
// instead of 64-bit vector address arithmetic.
// Walls and particles (kernels instantiated with EXT = true).  Node types are
// bit flags of the masks of the reference's set_walls (S:326-395).
enum : unsigned { NT_SOLID = 1u, NT_POUT = 2u, NT_NEAR = 4u, NT_REFILL = 8u, NT_PIN = 16u };
struct Ext {
    int wall[3][2];          // 1: half-way bounce-back wall face with a ghost layer (S:1557-1590)
    int mirror[3][2];        // 1: specular mirror face with a ghost layer (S:1443-1542)
    int npar;
    int pad;
    const unsigned *ntype;   // node types of the masks in force
    const int *pout;         // live-fluid particle id, 1-based (S:365)
    const double *center;    // particle centres [npar][3], 1-based lattice coordinates
    const double *cita;      // contact angle per particle (S:21)
    double *mac;             // 6 x N: rho, rho_R, rho_B, ux, uy, uz at NEAR nodes
    double *der;             // 6 x N: gradx, grady, gradz, n_x, n_y, n_z at NEAR nodes
    // factorised layout: red/blue side arrays (19 x N each, RB storage) of
    // the input and output states, which hold the slots of nodes that were
    // solid when the state was written
    const void *rbi[2];
    void *rbo[2];
};

// Recolour scalars of a factorised node that was solid when its step wrote
// the state: the colour fraction slot holds this value (a fraction never
// exceeds 1), and the node's slots are in the red/blue side arrays.
constexpr float FS_SOLID = 4.0f;

template <class T> __device__ __forceinline__ const T &at(const T *base, unsigned idx)
{
    return *(const T *)((const char *)base + idx * (unsigned)sizeof(T));
}
template <class T> __device__ __forceinline__ T &at(T *base, unsigned idx)
{
    return *(T *)((char *)base + idx * (unsigned)sizeof(T));
}

// Direction arrays in groups of four: direction i lives in the group that
// starts at base + (i & ~3) N, at element (i & 3) N + k.  Four arrays span
// at most 4 GiB for every box cgl_create accepts (N <= 2^28 for values of
// 4 bytes or less, N <= 2^27 for 8-byte ones), so a kernel holds five 64-bit
// bases per state array instead of nineteen (for a compile-time i the group
// bases are shared), which keeps them out of scalar-register spills.
template <class T> __device__ __forceinline__ T *grp(T *base, int i, size_t N)
{
    return base + (size_t)(i & ~3) * N;
}
__device__ __forceinline__ unsigned grp_off(int i, size_t N, unsigned k)
{
    return (unsigned)(i & 3) * (unsigned)N + k;
}

template <class S> struct StoreOf;
template <> struct StoreOf<double> {
    __device__ static double ld(const double *p) { return *p; }
    __device__ static void st(double *p, double v) { *p = v; }
    __device__ static double ld(const double *p, unsigned i) { return at(p, i); }
    __device__ static void st(double *p, unsigned i, double v) { at(p, i) = v; }
};
template <> struct StoreOf<float> {
    __device__ static float ld(const float *p) { return *p; }
    __device__ static void st(float *p, float v) { *p = v; }
    __device__ static float ld(const float *p, unsigned i) { return at(p, i); }
    __device__ static void st(float *p, unsigned i, float v) { at(p, i) = v; }
};
template <> struct StoreOf<_Float16> {
    __device__ static float ld(const _Float16 *p) { return (float)*p; }
    __device__ static void st(_Float16 *p, float v) { *p = (_Float16)v; }
    __device__ static float ld(const _Float16 *p, unsigned i) { return (float)at(p, i); }
    __device__ static void st(_Float16 *p, unsigned i, float v) { at(p, i) = (_Float16)v; }
};

// Arithmetic helpers: FP64 keeps IEEE division and square root (the FP64 path
// reproduces the CPU path bit for bit); FP32 uses the hardware reciprocal
// (1 ulp), multiplies by compile-time reciprocals of constants, and rsqrt.
__device__ __forceinline__ double fdiv(double a, double b) { return a / b; }
__device__ __forceinline__ float fdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double fsqrt(double v) { return __builtin_sqrt(v); }
__device__ __forceinline__ float fsqrt(float v) { return __builtin_sqrtf(v); }
template <class R> __device__ __forceinline__ R div_c(R a, double c)
{
    if constexpr (sizeof(R) == 8)
        return a / R(c);
    else
        return a * R(1.0 / c);
}

__device__ __forceinline__ int wrap1(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }
__device__ __forceinline__ int wrapm(int v, int n)
{
    v %= n;
    return v < 0 ? v + n : v;
}
// The same for the particle kernels, whose coordinates are nearly always
// within one period of the box: a compare and a select, and the division
// (tens of instructions for a run-time n) only for lanes beyond that.
__device__ __forceinline__ int wrapn(int v, int n)
{
    const int w = v < 0 ? v + n : (v >= n ? v - n : v);
    if ((unsigned)w < (unsigned)n) return w;
    return wrapm(v, n);
}

// Workgroups go to the 8 XCDs of the MI355X round-robin by linear id, and each
// XCD has its own L2.  Remap the linear id so that every XCD gets one
// contiguous range of logical tiles: neighbouring tiles, whose halo cache
// lines overlap, then share an L2 instead of fetching those lines twice.
constexpr unsigned NXCD = 8;
__device__ __forceinline__ void tile_of_block(const Lat &L, unsigned &bx, unsigned &by, unsigned &bz)
{
    const unsigned gx = gridDim.x, gy = gridDim.y, nb = gx * gy * gridDim.z;
    unsigned lin = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    if (L.xcd_remap == 2) {
        // whole z columns of tiles (gx tiles sharing misaligned halo lines) on
        // one XCD, consecutive columns dealt out over the XCDs: XCD k's j-th
        // block takes column (j / gx) * 8 + k, tile j % gx of it
        const unsigned span = NXCD * gx, fullg = nb - nb % span;
        if (lin < fullg) {
            const unsigned k = lin % NXCD, j = lin / NXCD;
            lin = ((j / gx) * NXCD + k) * gx + j % gx;
        }
    } else if (L.xcd_remap) {
        const unsigned full = nb - nb % NXCD;
        if (lin < full) lin = (lin % NXCD) * (full / NXCD) + lin / NXCD;
    }
    bx = lin % gx;
    by = (lin / gx) % gy;
    bz = lin / (gx * gy);
}

// ------------------------------------------------------------ state layouts
// RB  red and blue populations, 2 x 19 SoA arrays (the reference's ri, bi
//     right before streaming, S:249-250).
// FS  factorised red/blue: the post-collision total f_i (19 SoA arrays) and
//     per-node recolour scalars such that the node's red populations are
//       r_i = A f_i + w_i (G . c_i),   b_i = f_i - r_i,
//     which is exactly the form Recolor produces (S:713-726), so the factored
//     state holds the same red/blue populations in 19 + (4 or 6) values.
//     FP64 keeps A, gamma and n separately, G = gamma n, so a reader rebuilds
//     r_i with the reference's own rounding: r = A*f + (gamma*w_i)*(n . c_i).
enum { RB = 0, FS = 1 };
template <bool SHIFT> struct FsScal;
template <> struct FsScal<false> { static constexpr int NS = 6; }; // A, gamma, nx, ny, nz, unused (FP64)
template <> struct FsScal<true> { static constexpr int NS = 4; };  // A, Gx, Gy, Gz (FP32)

// n . c_i with zero terms skipped, in the reference's order (S:717).
template <class R> __device__ __forceinline__ R dot_c(int i, R nx, R ny, R nz)
{
    R dot = 0;
    bool first = true;
    if (CX[i] != 0) { dot = CX[i] == 1 ? nx : -nx; first = false; }
    if (CY[i] != 0) { R t = CY[i] == 1 ? ny : -ny; dot = first ? t : dot + t; first = false; }
    if (CZ[i] != 0) { R t = CZ[i] == 1 ? nz : -nz; dot = first ? t : dot + t; }
    return dot;
}

// Gathered node values.
//  FP64: r, b = the reference's post-stream red and blue populations.
//  shifted: ft_i = f_i - w_i rho0 (deviation of the total), sr = rho_R and
//  sb = rho_B, each summed directly (never as the difference rho - rho_R, which
//  would leave the minority colour of a nearly pure node to FP32 rounding).
template <class R, bool SHIFT> struct Node;
template <class R> struct Node<R, false> {
    R r[Q], b[Q];
};
template <class R> struct Node<R, true> {
    R ft[Q];
    R sr, sb;
};

// Shifted storage (FP32/FP16) holds each colour's populations as deviations
// from w_i rho0 s of the node that reads them, with the red and blue shift
// fractions s_red + s_blue = 1 taken from that node's phase field: the
// minority colour's fraction hm = (1 - |phi|)/2 (exact: 1 - |phi| has no
// rounding for |phi| >= 1/2) and the majority's 1 - hm.  A colour that is
// nearly absent at a node thus keeps FP32's relative precision through the
// shifts; pure phases shift by exactly 0 and 1.
template <class R> struct Shift {
    R hm;        // minority fraction
    bool redmaj; // phi >= 0: red is the majority colour
};
template <class R, class P> __device__ __forceinline__ Shift<R> shift_of(P ph)
{
    Shift<R> s;
    s.redmaj = ph >= P(0);
    s.hm = R(0.5) * (R(1) - R(ph < P(0) ? -ph : ph));
    return s;
}
// Red shift fraction of a node with shift sh minus that of a node with shift
// t, with the rounding of a difference of two small fractions where both
// nodes have the same majority colour.
template <class R> __device__ __forceinline__ R red_shift_diff(const Shift<R> &sh, const Shift<R> &t)
{
    if (sh.redmaj == t.redmaj) return sh.redmaj ? t.hm - sh.hm : sh.hm - t.hm;
    return sh.redmaj ? (R(1) - sh.hm) - t.hm : sh.hm - (R(1) - t.hm);
}

// Shift fractions in FP64 (exact for an FP32 phase field): host-layout
// conversions and the particle kernels.
__device__ __forceinline__ double shift_d(float ph, int red)
{
    const double hm = 0.5 * (1.0 - (double)(ph < 0.0f ? -ph : ph));
    const double sred = ph >= 0.0f ? 1.0 - hm : hm;
    return red ? sred : 1.0 - sred;
}

// Signed colour fraction of the factorised layout's FP32 recolour scalars:
// +A (red fraction) where blue is the majority, -B (blue fraction) where red
// is, so that the minority fraction is stored to FP32's relative precision;
// a pure red node stores -0, a pure blue one +0.
template <class R> __device__ __forceinline__ void decode_frac(float x, R &A, R &B);

// The scatter form's phase-field buffer holds the signed minority fraction q
// of each node's colour split: -rho_B/rho where red is the majority (-0 for a
// pure red node), +rho_R/rho where blue is (+0 for pure blue), so that a
// nearly absent colour keeps FP32's relative precision; phi = 1 + 2q or
// 2q - 1.
template <class R> __device__ __forceinline__ float minority_q(R sr, R sb, R irho)
{
    if (sb <= sr) return -__builtin_fmaxf(float(sb * irho), 0.0f);
    return __builtin_fmaxf(float(sr * irho), 0.0f);
}
__device__ __forceinline__ float phi_of_q(float q) { return __builtin_signbitf(q) ? 1.0f + 2.0f * q : 2.0f * q - 1.0f; }

// Fixed-point units of the scatter form's phase-field sums: 2^29 per rho0.
constexpr float SC_ONE = 536870912.0f;

// floor(v + 0.5) as an int in one instruction; its result equals the exact
// floor(v + 0.5) on gfx950 for every value tried (1M floats: |v| < 2^30,
// exact halves, integers above 2^23 where v + 0.5 is not a float).
__device__ __forceinline__ int round_half_up(float v)
{
    int r;
    asm("v_cvt_rpi_i32_f32 %0, %1" : "=v"(r) : "v"(v));
    return r;
}

// An entry of that buffer seen as phi: it converts from phi on writes and to
// phi on reads, so that the wall, particle and band kernels (templates on the
// phase-field type) run on the scatter form's buffer.  raw() stores q itself.
struct QPhi {
    float q;
    QPhi() = default;
    __device__ QPhi(float phi) : q(phi >= 0.0f ? -0.5f * (1.0f - phi) : 0.5f * (1.0f + phi)) {}
    __device__ operator float() const { return phi_of_q(q); }
    __device__ static QPhi raw(float v)
    {
        QPhi e;
        e.q = v;
        return e;
    }
};
static_assert(sizeof(QPhi) == sizeof(float), "QPhi is a float in memory");

template <class R> __device__ __forceinline__ void decode_frac(float x, R &A, R &B)
{
    if (__builtin_signbitf(x)) {
        B = R(-x);
        A = R(1) - B;
    } else {
        A = R(x);
        B = R(1) - A;
    }
}

// Recolour scalars of one node in the FP64 factorised layout.
struct FsScal64 {
    double a, g, nx, ny, nz;
};

// Pull gather of the factorised layout: f_i from x - c_i and the red part
// rebuilt with the source node's scalars, which `scal(i, m)` returns for
// direction i and source node m (from global memory or from an LDS copy).
//  FP64:    r_i = A f_i + (gamma w_i)(n . c_i), the reference's rounding.
//  shifted: rho_R = rho0 sum_i w_i A_i + sum_i A_i f~_i + sum_i w_i G_i.c_i:
//           the O(1) part is summed per weight class, the rest are small
//           deviations.
template <class S, class R, bool SHIFT, class Scal>
__device__ __forceinline__ void gather_fs(const S *__restrict__ fs, size_t N, const unsigned xo[3],
                                          const unsigned yo[3], const unsigned zo[3], Scal scal,
                                          Node<R, SHIFT> &nd, R rho0)
{
    if constexpr (!SHIFT) {
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            const unsigned m = xo[1 - CX[i]] + yo[1 - CY[i]] + zo[1 - CZ[i]];
            R f = StoreOf<S>::ld(grp(fs, i, N), grp_off(i, N, m));
            const FsScal64 sc = scal(i, m);
            R r = sc.a * f;
            if (i > 0) r = r + (sc.g * R(W[i])) * dot_c<R>(i, sc.nx, sc.ny, sc.nz);
            nd.r[i] = r;
            nd.b[i] = f - r;
        }
    } else {
        R a0 = 0, a1 = 0, a2 = 0, b0 = 0, b1 = 0, b2 = 0, dva = 0, dvb = 0, g1 = 0, g2 = 0;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            const unsigned m = xo[1 - CX[i]] + yo[1 - CY[i]] + zo[1 - CZ[i]];
            R f = StoreOf<S>::ld(grp(fs, i, N), grp_off(i, N, m));
            const float4 g = scal(i, m);
            R A, B;
            decode_frac(g.x, A, B);
            dva += A * f;
            dvb += B * f;
            R gc = 0;
            if (CX[i] == 1) gc += R(g.y);
            if (CX[i] == -1) gc -= R(g.y);
            if (CY[i] == 1) gc += R(g.z);
            if (CY[i] == -1) gc -= R(g.z);
            if (CZ[i] == 1) gc += R(g.w);
            if (CZ[i] == -1) gc -= R(g.w);
            if (i == 0) { a0 = A; b0 = B; }
            else if (i <= 6) { a1 += A; b1 += B; g1 += gc; }
            else { a2 += A; b2 += B; g2 += gc; }
            nd.ft[i] = f;
        }
        const R gw = g1 * R(W[1]) + g2 * R(W[7]);
        nd.sr = rho0 * (a0 * R(W[0]) + a1 * R(W[1]) + a2 * R(W[7])) + dva + gw;
        nd.sb = rho0 * (b0 * R(W[0]) + b1 * R(W[1]) + b2 * R(W[7])) + dvb - gw;
    }
}

// Pull gather at node (x, y, z) from x - c_i (streaming, S:743-753).
// ax: this node's phase field of the stored state, which fixes its shift
// fractions (RB shifted storage only).
// Half-way wall bounce-back in pull form: a population whose upstream node is
// a wall ghost layer is the node's own opposite population (boundary then
// streaming, S:1557-1590 and S:743-753).  bbmask bit i set = direction i.
struct WallBB {
    unsigned mask = 0;
};
__device__ __forceinline__ bool is_wall_ghost(const Ext &E, const Lat &L, int x, int y, int z)
{
    return (E.wall[0][0] && x == 0) || (E.wall[0][1] && x == L.nx - 1) || (E.wall[1][0] && y == 0) ||
           (E.wall[1][1] && y == L.ny - 1) || (E.wall[2][0] && z == 0) || (E.wall[2][1] && z == L.nz - 1);
}
// Mirror ghost planes crossed by (x, y, z): bit a set for axis a.
__device__ __forceinline__ unsigned mirror_axes(const Ext &E, const Lat &L, int x, int y, int z)
{
    const int c[3] = {x, y, z}, n[3] = {L.nx, L.ny, L.nz};
    unsigned A = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a)
        if ((E.mirror[a][0] && c[a] == 0) || (E.mirror[a][1] && c[a] == n[a] - 1)) A |= 1u << a;
    return A;
}
// Ghost layer of a wall or a mirror face: not a lattice node of the fluid.
__device__ __forceinline__ bool is_ghost(const Ext &E, const Lat &L, int x, int y, int z)
{
    return is_wall_ghost(E, L, x, y, z) || mirror_axes(E, L, x, y, z) != 0;
}
// Mirror ghost layers next to a live node: bit 2a (low side) and 2a+1 (high
// side) of axis a.
__device__ __forceinline__ unsigned mirror_adjacent(const Ext &E, const Lat &L, int x, int y, int z)
{
    const int c[3] = {x, y, z}, n[3] = {L.nx, L.ny, L.nz};
    unsigned m = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (E.mirror[a][0] && c[a] == 1) m |= 1u << (2 * a);
        if (E.mirror[a][1] && c[a] == n[a] - 2) m |= 1u << (2 * a + 1);
    }
    return m;
}
// Axes along which the upstream node x - c_i of direction i lies in a mirror
// ghost layer, from the adjacency bits of the node.
template <int I> __device__ __forceinline__ unsigned mirrored_axes(unsigned madj)
{
    unsigned A = 0;
    if ((CX[I] == 1 && (madj & 1u)) || (CX[I] == -1 && (madj & 2u))) A |= 1u;
    if ((CY[I] == 1 && (madj & 4u)) || (CY[I] == -1 && (madj & 8u))) A |= 2u;
    if ((CZ[I] == 1 && (madj & 16u)) || (CZ[I] == -1 && (madj & 32u))) A |= 4u;
    return A;
}
// The directions of wall_bb_mask that come from the walls of one axis a
// (coordinate c of n along it): a march along x holds those of y and z.
__device__ __forceinline__ unsigned wall_bb_axis(const Ext &E, int a, int c, int n)
{
    const bool lo = E.wall[a][0] && c == 1, hi = E.wall[a][1] && c == n - 2;
    unsigned m = 0;
#pragma unroll
    for (int i = 1; i < Q; ++i) {
        const int ci = a == 0 ? CX[i] : (a == 1 ? CY[i] : CZ[i]);
        if ((ci == 1 && lo) || (ci == -1 && hi)) m |= 1u << i;
    }
    return m;
}
__device__ __forceinline__ unsigned wall_bb_mask(const Ext &E, const Lat &L, int x, int y, int z)
{
    const bool lo[3] = {E.wall[0][0] && x == 1, E.wall[1][0] && y == 1, E.wall[2][0] && z == 1};
    const bool hi[3] = {E.wall[0][1] && x == L.nx - 2, E.wall[1][1] && y == L.ny - 2, E.wall[2][1] && z == L.nz - 2};
    unsigned m = 0;
#pragma unroll
    for (int i = 1; i < Q; ++i) {
        // the upstream node x - c_i lies in a ghost layer
        const bool g = (CX[i] == 1 && lo[0]) || (CX[i] == -1 && hi[0]) || (CY[i] == 1 && lo[1]) ||
                       (CY[i] == -1 && hi[1]) || (CZ[i] == 1 && lo[2]) || (CZ[i] == -1 && hi[2]);
        if (g) m |= 1u << i;
    }
    return m;
}

// Lanes that the tile kernels of the factorised layout with walls, mirrors or
// particles leave to the per-node kernels (k_moments_special,
// k_collide_special): ghost layers, nodes next to a mirror ghost layer, and
// the particles' band (NEAR; it contains every solid and refilled node and
// every node next to one).  Nodes next to a wall stay in the tile kernels
// (rev_scal).
__device__ __forceinline__ bool special_lane(const Ext &E, const Lat &L, int x, int y, int z, unsigned n)
{
    return is_ghost(E, L, x, y, z) || mirror_adjacent(E, L, x, y, z) || (E.ntype && (at(E.ntype, n) & NT_NEAR));
}

// Source slot of direction i at a live node in the pull layout: the upstream
// node x - c_i, or with a wall ghost there the node's own opposite slot
// (bounce-back, bb), or with mirror ghosts there (madj) the slot of the
// mirrored direction at the upstream node's mirror image, which is what the
// reference's ghost fill followed by streaming delivers (S:1443-1542,
// S:735-768).  Walls take precedence: their bounce-back is written last
// (S:1557-1590).
template <int I>
__device__ __forceinline__ void pull_slot(const unsigned xo[3], const unsigned yo[3], const unsigned zo[3],
                                          unsigned bb, unsigned madj, unsigned &m, int &d)
{
    if ((bb >> I) & 1u) {
        m = xo[1] + yo[1] + zo[1];
        d = OPC[I];
        return;
    }
    const unsigned A = mirrored_axes<I>(madj);
    // the upstream slot, moved back onto the node's own coordinate along the
    // mirrored axes (differences of constant-index offsets, so the offset
    // arrays stay in registers)
    const unsigned dx = xo[1] - xo[1 - CX[I]], dy = yo[1] - yo[1 - CY[I]], dz = zo[1] - zo[1 - CZ[I]];
    m = xo[1 - CX[I]] + yo[1 - CY[I]] + zo[1 - CZ[I]];
    m += ((A & 1u) ? dx : 0u) + ((A & 2u) ? dy : 0u) + ((A & 4u) ? dz : 0u);
    // the reflected direction by selects over the seven masks (a runtime
    // index into the table would place a copy of its column in scratch)
    d = I;
#pragma unroll
    for (unsigned a = 1; a < 8; ++a)
        if (A == a) d = REFL.d[a][I];
}

template <int LIN, class S, class R, bool SHIFT>
__device__ __forceinline__ void gather(const Lat &L, const void *p0, const void *p1, const unsigned xo[3],
                                       const unsigned yo[3], const unsigned zo[3], R ax, Node<R, SHIFT> &nd,
                                       unsigned bb = 0, unsigned madj = 0)
{
    const size_t N = (size_t)L.N;
    const R rho0 = R(L.rho0);
    if constexpr (LIN == RB) {
        const S *rs = (const S *)p0, *bs = (const S *)p1;
        if constexpr (!SHIFT) {
            static_for<Q>([&](auto ic) {
                constexpr int i = decltype(ic)::value;
                unsigned m;
                int d;
                pull_slot<i>(xo, yo, zo, bb, madj, m, d);
                const unsigned k = grp_off(d, N, m);
                nd.r[i] = StoreOf<S>::ld(grp(rs, d, N), k);
                nd.b[i] = StoreOf<S>::ld(grp(bs, d, N), k);
            });
        } else {
            // deviations from w_i rho0 s_red and w_i rho0 s_blue of this node:
            // small, summed plainly, each colour by itself
            R sr = 0, sb = 0;
            static_for<Q>([&](auto ic) {
                constexpr int i = decltype(ic)::value;
                unsigned m;
                int d;
                pull_slot<i>(xo, yo, zo, bb, madj, m, d);
                const unsigned k = grp_off(d, N, m);
                R rv = StoreOf<S>::ld(grp(rs, d, N), k);
                R bv = StoreOf<S>::ld(grp(bs, d, N), k);
                sr += rv;
                sb += bv;
                nd.ft[i] = rv + bv;
            });
            const Shift<R> sh = shift_of<R>(ax);
            const R smaj = R(1) - sh.hm;
            nd.sr = sr + rho0 * (sh.redmaj ? smaj : sh.hm);
            nd.sb = sb + rho0 * (sh.redmaj ? sh.hm : smaj);
        }
    } else {
        const S *fs = (const S *)p0;
        if constexpr (!SHIFT) {
            const double *sc = (const double *)p1;
            gather_fs<S, R, SHIFT>(
                fs, N, xo, yo, zo,
                [&](int, unsigned m) {
                    const double2 s01 = at((const double2 *)sc, 3 * m);
                    const double2 s23 = at((const double2 *)sc, 3 * m + 1);
                    return FsScal64{s01.x, s01.y, s23.x, s23.y, at(sc, 6 * m + 4)};
                },
                nd, rho0);
        } else {
            const float4 *sc = (const float4 *)p1;
            gather_fs<S, R, SHIFT>(fs, N, xo, yo, zo, [&](int, unsigned m) { return at(sc, m); }, nd, rho0);
        }
    }
}

// The factorised layout with walls, mirrors and particles keeps every pull
// direction-uniform: slot i of node x is read from node x - c_i, as in the
// periodic gather.  After each step k_fill_ghosts copies into the wall and
// mirror ghost slots the totals that the reference's boundary fills put
// there (S:1443-1592): the reader's own opposite slot off a wall, the mirror
// image's reflected slot across mirror faces.  The reader splits such a slot
// with the recolour scalars of the node it came from, its own or the mirror
// image's, in that slot's direction: c_i reversed on all axes or on the
// mirrored ones.  fs_src gives that node's offset o from x (-1, 0, +1 per
// axis, an index into the LDS rings) and the reversed axes (flip).
template <int I> __device__ __forceinline__ void fs_src(unsigned bb, unsigned madj, int o[3], unsigned &flip)
{
    if ((bb >> I) & 1u) {
        flip = 7u;
        o[0] = o[1] = o[2] = 0;
        return;
    }
    flip = mirrored_axes<I>(madj);
    o[0] = (flip & 1u) ? 0 : -CX[I];
    o[1] = (flip & 2u) ? 0 : -CY[I];
    o[2] = (flip & 4u) ? 0 : -CZ[I];
}

// dot_c of the direction c_I with the components of the axes in `flip`
// reversed: the same terms in the same order with exact sign changes, so it
// rounds as dot_c of the reflected direction.
template <int I, class R> __device__ __forceinline__ R dot_cf(unsigned flip, R nx, R ny, R nz)
{
    R dot = 0;
    bool first = true;
    if (CX[I] != 0) { R t = CX[I] == 1 ? nx : -nx; if (flip & 1u) t = -t; dot = t; first = false; }
    if (CY[I] != 0) { R t = CY[I] == 1 ? ny : -ny; if (flip & 2u) t = -t; dot = first ? t : dot + t; first = false; }
    if (CZ[I] != 0) { R t = CZ[I] == 1 ? nz : -nz; if (flip & 4u) t = -t; dot = first ? t : dot + t; }
    return dot;
}
// G . c_I with the components of the axes in `flip` reversed (shifted storage).
template <int I, class R> __device__ __forceinline__ R gdot_f(unsigned flip, float gx, float gy, float gz)
{
    R gc = 0;
    if (CX[I] != 0) { R t = CX[I] == 1 ? R(gx) : -R(gx); gc += (flip & 1u) ? -t : t; }
    if (CY[I] != 0) { R t = CY[I] == 1 ? R(gy) : -R(gy); gc += (flip & 2u) ? -t : t; }
    if (CZ[I] != 0) { R t = CZ[I] == 1 ? R(gz) : -R(gz); gc += (flip & 4u) ? -t : t; }
    return gc;
}

// Nodes next to a wall in the tile kernels of the factorised layout: a slot
// read from a wall ghost (bit i of bb) holds the node's own opposite slot
// (k_fill_ghosts) and is split with the node's own scalars in the reversed
// direction, i.e. with its gradient term negated (rev_scal): the same
// rounding as dot_c of the opposite direction.
__device__ __forceinline__ float4 rev_scal(const float4 &g) { return make_float4(g.x, -g.y, -g.z, -g.w); }
__device__ __forceinline__ FsScal64 rev_scal(const FsScal64 &g) { return FsScal64{g.a, g.g, -g.nx, -g.ny, -g.nz}; }

// Recolour scalars of node m in the FP64 factorised layout.
__device__ __forceinline__ FsScal64 fs64_at(const double *sc, unsigned m)
{
    const double2 s01 = at((const double2 *)sc, 3 * m);
    const double2 s23 = at((const double2 *)sc, 3 * m + 1);
    return FsScal64{s01.x, s01.y, s23.x, s23.y, at(sc, 6 * m + 4)};
}
__device__ __forceinline__ bool solid_scal(const FsScal64 &s) { return s.a > 0.5 * double(FS_SOLID); }
__device__ __forceinline__ bool solid_scal(const float4 &s) { return s.x > 0.5f * FS_SOLID; }

// Slots of a node that was solid when its step wrote the state (FS_SOLID
// scalars) are no colour split of one node: the populations streaming through
// a solid node (S:622, S:712) and the moving-wall bounce-back values of the
// curve links (S:1004-1018).  The total is in the population array as for any
// node; the colours are in side arrays: FP64 the red and blue values (r, b),
// shifted storage the slot's own minority fraction (FP32), with which the
// reader splits the total as it splits a factorised slot with no gradient
// term.  A slot's minority population can be negative (recolouring leaves
// w_i beta rho_R rho_B / rho (n . c_i) on it), so the fraction keeps its
// sign, and the lowest mantissa bit says which colour it belongs to (1: blue,
// red being the majority).
template <class R> __device__ __forceinline__ float enc_slot(R r, R b)
{
    const bool redmaj = b <= r;
    const float q = float((redmaj ? b : r) / (r + b));
    return __uint_as_float((__float_as_uint(q) & ~1u) | (redmaj ? 1u : 0u));
}
template <class R> __device__ __forceinline__ void dec_slot(float v, R &A, R &B)
{
    const unsigned u = __float_as_uint(v);
    const R q = R(__uint_as_float(u & ~1u));
    if (u & 1u) {
        B = q;
        A = R(1) - q;
    } else {
        A = q;
        B = R(1) - q;
    }
}

// Pull gather of the factorised layout with walls, mirrors and particles:
// f_i from x - c_i, split with the scalars that scal(o) returns for the node
// at offset o (fs_src).  SOLIDS (nodes in the particles' band): where those
// scalars are FS_SOLID, the colours come from the side arrays instead (a
// node outside the band has no solid node within reach).
template <bool SOLIDS, class S, class R, bool SHIFT, class Scal>
__device__ __forceinline__ void gather_fs_ext(const S *__restrict__ fs, const void *__restrict__ side0,
                                              const void *__restrict__ side1, size_t N, const unsigned xo[3],
                                              const unsigned yo[3], const unsigned zo[3], unsigned bb,
                                              unsigned madj, Scal scal, Node<R, SHIFT> &nd, R rho0)
{
    if constexpr (!SHIFT) {
        static_for<Q>([&](auto ic) {
            constexpr int i = decltype(ic)::value;
            const unsigned m = xo[1 - CX[i]] + yo[1 - CY[i]] + zo[1 - CZ[i]], k = grp_off(i, N, m);
            int o[3];
            unsigned flip;
            fs_src<i>(bb, madj, o, flip);
            const FsScal64 sc = scal(o);
            if (SOLIDS && solid_scal(sc)) {
                nd.r[i] = StoreOf<S>::ld(grp((const S *)side0, i, N), k);
                nd.b[i] = StoreOf<S>::ld(grp((const S *)side1, i, N), k);
            } else {
                const R f = StoreOf<S>::ld(grp(fs, i, N), k);
                R r = sc.a * f;
                if (i > 0) r = r + (sc.g * R(W[i])) * dot_cf<i, R>(flip, sc.nx, sc.ny, sc.nz);
                nd.r[i] = r;
                nd.b[i] = f - r;
            }
        });
    } else {
        R a0 = 0, a1 = 0, a2 = 0, b0 = 0, b1 = 0, b2 = 0, dva = 0, dvb = 0, g1 = 0, g2 = 0;
        static_for<Q>([&](auto ic) {
            constexpr int i = decltype(ic)::value;
            const unsigned m = xo[1 - CX[i]] + yo[1 - CY[i]] + zo[1 - CZ[i]], k = grp_off(i, N, m);
            int o[3];
            unsigned flip;
            fs_src<i>(bb, madj, o, flip);
            const R f = StoreOf<S>::ld(grp(fs, i, N), k);
            float4 g = scal(o);
            R A, B;
            if (SOLIDS && solid_scal(g)) {
                dec_slot(at(grp((const float *)side0, i, N), k), A, B);
                g.y = g.z = g.w = 0.f;
            } else {
                decode_frac(g.x, A, B);
            }
            dva += A * f;
            dvb += B * f;
            const R gc = gdot_f<i, R>(flip, g.y, g.z, g.w);
            if (i == 0) { a0 = A; b0 = B; }
            else if (i <= 6) { a1 += A; b1 += B; g1 += gc; }
            else { a2 += A; b2 += B; g2 += gc; }
            nd.ft[i] = f;
        });
        const R gw = g1 * R(W[1]) + g2 * R(W[7]);
        nd.sr = rho0 * (a0 * R(W[0]) + a1 * R(W[1]) + a2 * R(W[7])) + dva + gw;
        nd.sb = rho0 * (b0 * R(W[0]) + b1 * R(W[1]) + b2 * R(W[7])) + dvb - gw;
    }
}

// Colours of one slot (total f; deviation f~ in shifted storage) split with
// the scalars sc of its node in direction c_I reversed on the axes in flip:
// FP64 the red and blue values with the reference's rounding, shifted
// storage the slot's encoded minority fraction (the minority formed first).
// FS_SOLID scalars: the side arrays' colours of slot k.
template <int I, class S, class R, bool SHIFT, class Sc>
__device__ __forceinline__ void slot_side(const void *__restrict__ side0, const void *__restrict__ side1, size_t N,
                                          unsigned k, R f, const Sc &sc, unsigned flip, R rho0, R &vr, R &vb,
                                          float &q)
{
    if constexpr (!SHIFT) {
        (void)rho0;
        (void)q;
        if (solid_scal(sc)) {
            vr = StoreOf<S>::ld(grp((const S *)side0, I, N), k);
            vb = StoreOf<S>::ld(grp((const S *)side1, I, N), k);
        } else {
            R r = sc.a * f;
            if (I > 0) r = r + (sc.g * R(W[I])) * dot_cf<I, R>(flip, sc.nx, sc.ny, sc.nz);
            vr = r;
            vb = f - r;
        }
    } else {
        (void)side1;
        (void)vr;
        (void)vb;
        if (solid_scal(sc)) {
            q = at(grp((const float *)side0, I, N), k);
        } else {
            const R fa = R(W[I]) * rho0 + f, gd = R(W[I]) * gdot_f<I, R>(flip, sc.y, sc.z, sc.w);
            R A, B;
            decode_frac(sc.x, A, B);
            R r, b;
            if (__builtin_signbitf(sc.x)) {
                b = B * fa - gd;
                r = fa - b;
            } else {
                r = A * fa + gd;
                b = fa - r;
            }
            q = enc_slot(r, b);
        }
    }
}

// Moments of the gathered populations.
//  FP64: calmacros (S:792-799) operation for operation.
//  shifted: compensated sums of the deviations.
template <class R, bool SHIFT> struct Moments {
    R sr, sb, rho, ux, uy, uz, phi;
    R drho; // shifted storage: the sum of the deviations, rho - rho0
};

template <class R, bool SHIFT>
__device__ __forceinline__ Moments<R, SHIFT> moments(const Lat &L, const Node<R, SHIFT> &nd)
{
    Moments<R, SHIFT> m;
    if constexpr (!SHIFT) {
        R f[Q];
#pragma unroll
        for (int i = 0; i < Q; ++i) f[i] = nd.r[i] + nd.b[i];
        R sr = 0, sb = 0;
#pragma unroll
        for (int i = 0; i < Q; ++i) sr += nd.r[i];
#pragma unroll
        for (int i = 0; i < Q; ++i) sb += nd.b[i];
        R jx = 0, jy = 0, jz = 0;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            if (CX[i] == 1) jx += f[i];
            if (CX[i] == -1) jx -= f[i];
        }
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            if (CY[i] == 1) jy += f[i];
            if (CY[i] == -1) jy -= f[i];
        }
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            if (CZ[i] == 1) jz += f[i];
            if (CZ[i] == -1) jz -= f[i];
        }
        m.sr = sr;
        m.sb = sb;
        m.rho = sr + sb;
        m.ux = jx / m.rho;
        m.uy = jy / m.rho;
        m.uz = jz / m.rho;
        m.phi = (sr - sb) / m.rho;  // FP64 path: IEEE division as the reference
    } else {
        // the density deviation is the sum of the stored deviations: the O(1)
        // part rho0 is never summed, so no compensation is needed
        R df = 0, jx = 0, jy = 0, jz = 0;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            const R f = nd.ft[i];
            df += f;
            if (CX[i] == 1) jx += f;
            if (CX[i] == -1) jx -= f;
            if (CY[i] == 1) jy += f;
            if (CY[i] == -1) jy -= f;
            if (CZ[i] == 1) jz += f;
            if (CZ[i] == -1) jz -= f;
        }
        m.rho = R(L.rho0) + df;
        m.drho = df;
        m.sr = nd.sr;
        m.sb = nd.sb;
        const R irho = fdiv(R(1), m.rho);
        m.ux = jx * irho;
        m.uy = jy * irho;
        m.uz = jz * irho;
        // phi = (rho_R - rho_B)/rho from the minority colour (S:799)
        m.phi = m.sb <= m.sr ? R(1) - R(2) * m.sb * irho : R(2) * m.sr * irho - R(1);
    }
    return m;
}

// Node offsets (already multiplied by strides) for coordinates c-1, c, c+1.
__device__ __forceinline__ void node_offsets(const Lat &L, int x, int y, int z, unsigned xo[3], unsigned yo[3],
                                             unsigned zo[3])
{
    const unsigned sx = (unsigned)L.ny * (unsigned)L.nz;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        xo[k] = (unsigned)wrap1(x + k - 1, L.nx) * sx;
        yo[k] = (unsigned)wrap1(y + k - 1, L.ny) * (unsigned)L.nz;
        zo[k] = (unsigned)wrap1(z + k - 1, L.nz);
    }
}

// The phase-field buffer entry of a node from its moments: phi, or for the
// scatter form's buffer (QPhi) the signed minority fraction from the colour
// densities, to their relative precision.
template <class P, class R, bool SHIFT> __device__ __forceinline__ P phi_entry(const Moments<R, SHIFT> &mo)
{
    if constexpr (std::is_same_v<P, QPhi>)
        return QPhi::raw(minority_q(mo.sr, mo.sb, fdiv(R(1), mo.rho)));
    else
        return P(mo.phi);
}

// ---------------------------------------------------------------- K1 moments
// Writes phi of the step.  With FULL it also writes the macroscopic fields
// rho, rho_R, rho_B, ux, uy, uz, phi (bare velocity, as calmacros) as FP64.
// One node of tile (bx, by) of plane bz.
// Moments of live node (x, y, z) of the stored state as calmacros forms them
// (S:770-831), with its node index n and node type nt; false on wall and
// mirror ghost layers, where calmacros computes none.
template <int LIN, class S, class R, class P, bool SHIFT, bool EXT>
__device__ __forceinline__ bool node_moments(const Lat &L, const void *__restrict__ p0, const void *__restrict__ p1,
                                             const P *__restrict__ phio, const Ext &E, int x, int y, int z,
                                             unsigned &n, Moments<R, SHIFT> &mo, unsigned &nt)
{
    if constexpr (EXT)
        if (is_ghost(E, L, x, y, z)) return false;
    unsigned xo[3], yo[3], zo[3];
    node_offsets(L, x, y, z, xo, yo, zo);
    n = xo[1] + yo[1] + zo[1];
    R ax = 0; // the node's phase field of the stored state: its shift fractions
    if constexpr (SHIFT && LIN == RB) ax = R(at(phio, n));
    Node<R, SHIFT> nd;
    nt = 0;
    if constexpr (EXT) {
        const unsigned bb = wall_bb_mask(E, L, x, y, z), madj = mirror_adjacent(E, L, x, y, z);
        if constexpr (LIN == FS) {
            gather_fs_ext<true, S, R, SHIFT>(
                (const S *)p0, E.rbi[0], E.rbi[1], (size_t)L.N, xo, yo, zo, bb, madj,
                [&](const int o[3]) {
                    const unsigned m = xo[1 + o[0]] + yo[1 + o[1]] + zo[1 + o[2]];
                    if constexpr (SHIFT)
                        return at((const float4 *)p1, m);
                    else
                        return fs64_at((const double *)p1, m);
                },
                nd, R(L.rho0));
        } else {
            gather<LIN, S, R, SHIFT>(L, p0, p1, xo, yo, zo, ax, nd, bb, madj);
        }
        if (E.ntype) nt = at(E.ntype, n);
    } else {
        gather<LIN, S, R, SHIFT>(L, p0, p1, xo, yo, zo, ax, nd);
    }
    mo = moments<R, SHIFT>(L, nd);
    if constexpr (EXT) {
        if (nt & NT_SOLID) { // calmacros at solid nodes (S:784-789); rho_R, rho_B untouched
            mo.phi = 0;
            mo.rho = R(L.rho0);
            mo.ux = mo.uy = mo.uz = 0;
        }
    }
    return true;
}

// The macroscopic fields of a node in the particles' band (NEAR), which the
// particle kernels read (FP64): rho and the velocity everywhere in the band,
// the colour densities at fluid nodes (calmacros leaves them at solid nodes).
template <class R, bool SHIFT>
__device__ __forceinline__ void band_macros(const Lat &L, const Ext &E, unsigned n, unsigned nt,
                                            const Moments<R, SHIFT> &mo)
{
    if (nt & NT_NEAR) {
        const size_t N = (size_t)L.N;
        E.mac[n] = mo.rho;
        if (!(nt & NT_SOLID)) {
            E.mac[N + n] = mo.sr;
            E.mac[2 * N + n] = mo.sb;
        }
        E.mac[3 * N + n] = mo.ux;
        E.mac[4 * N + n] = mo.uy;
        E.mac[5 * N + n] = mo.uz;
    }
}

template <int LIN, class S, class R, class P, bool SHIFT, bool FULL, bool EXT>
__device__ __forceinline__ void moments_node(const Lat &L, const void *__restrict__ p0, const void *__restrict__ p1,
                                             const P *__restrict__ phio, P *__restrict__ phi,
                                             double *__restrict__ mac, const Ext &E, unsigned bx, unsigned by,
                                             unsigned bz)
{
    const int z = bx * TZ + threadIdx.x;
    const int y = by * TY + threadIdx.y;
    const int x = bz;
    if (z >= L.nz || y >= L.ny) return;
    unsigned n, nt;
    Moments<R, SHIFT> mo;
    if (!node_moments<LIN, S, R, P, SHIFT, EXT>(L, p0, p1, phio, E, x, y, z, n, mo, nt)) {
        // phi is copied from the adjacent layer afterwards (S:806); the
        // other fields are not computed on ghost layers and read as zero
        if constexpr (FULL) {
            const size_t N = (size_t)L.N;
            n = ((unsigned)x * (unsigned)L.ny + (unsigned)y) * (unsigned)L.nz + (unsigned)z;
#pragma unroll
            for (int k = 0; k < 6; ++k) mac[k * N + n] = 0.0;
        }
        return;
    }
    if constexpr (EXT) band_macros(L, E, n, nt, mo);
    at(phi, n) = P(mo.phi);
    if constexpr (FULL) {
        const size_t N = (size_t)L.N;
        mac[n] = mo.rho;
        mac[N + n] = mo.sr;
        mac[2 * N + n] = mo.sb;
        mac[3 * N + n] = mo.ux;
        mac[4 * N + n] = mo.uy;
        mac[5 * N + n] = mo.uz;
        mac[6 * N + n] = mo.phi;
    }
}

template <int LIN, class S, class R, class P, bool SHIFT, bool FULL, bool EXT = false>
__global__ __launch_bounds__(256) void k_moments(Lat L, const void *__restrict__ p0, const void *__restrict__ p1,
                                                 const P *__restrict__ phio, P *__restrict__ phi,
                                                 double *__restrict__ mac, Ext E = Ext{})
{
    unsigned bx, by, bz;
    tile_of_block(L, bx, by, bz);
    moments_node<LIN, S, R, P, SHIFT, FULL, EXT>(L, p0, p1, phio, phi, mac, E, bx, by, bz);
}

// K1 for the factorised layout, marching in x: the recolour scalars of
// planes x-1, x, x+1 (with a one-node y/z halo) sit in an LDS ring, so each
// node's scalars are fetched from memory once instead of by all 19 readers.
template <bool SHIFT, int ROWS = TY> struct FsxRing;
template <int ROWS> struct FsxRing<true, ROWS> {
    float4 s[3][ROWS + 2][TZ + 2];
};
template <int ROWS> struct FsxRing<false, ROWS> {
    double s[3][5][ROWS + 2][TZ + 2];
};

// ENC: write the signed minority fraction (the scatter form's phase-field
// buffer) instead of phi.  EXT (factorised layout with walls, mirrors or
// particles): the plain moments of every lane but the special ones
// (special_lane), which k_moments_special forms.
template <class S, class R, class P, bool SHIFT, bool ENC = false, bool EXT = false>
__device__ __forceinline__ void moments_fsx_tile(const Lat &L, const void *__restrict__ p0,
                                                 const void *__restrict__ p1, const P *__restrict__ phio,
                                                 P *__restrict__ phi, int xc, unsigned bx, unsigned by, unsigned bz,
                                                 const Ext &E = Ext{})
{
    __shared__ FsxRing<SHIFT> sm;
    const int tz = threadIdx.x, ty = threadIdx.y, tid = ty * TZ + tz;
    const int z0 = bx * TZ, y0 = by * TY, x0 = bz * xc;
    const int x1 = min(x0 + xc, L.nx);
    const int z = z0 + tz, y = y0 + ty;
    const bool active = z < L.nz && y < L.ny;
    const unsigned sx = (unsigned)L.ny * (unsigned)L.nz;
    const size_t N = (size_t)L.N;
    const S *fs = (const S *)p0;
    (void)phio;
    auto load = [&](int p) {
        const int sl = (p + 9) % 3;
        const unsigned gx = (unsigned)wrapm(p, L.nx) * sx;
        for (int e = tid; e < (TY + 2) * (TZ + 2); e += TZ * TY) {
            const int jy = e / (TZ + 2), jz = e - jy * (TZ + 2);
            const unsigned n = gx + (unsigned)wrapm(y0 - 1 + jy, L.ny) * (unsigned)L.nz + wrapm(z0 - 1 + jz, L.nz);
            if constexpr (SHIFT) {
                sm.s[sl][jy][jz] = at((const float4 *)p1, n);
            } else {
                const double *sc = (const double *)p1;
                const double2 s01 = at((const double2 *)sc, 3 * n);
                const double2 s23 = at((const double2 *)sc, 3 * n + 1);
                sm.s[sl][0][jy][jz] = s01.x;
                sm.s[sl][1][jy][jz] = s01.y;
                sm.s[sl][2][jy][jz] = s23.x;
                sm.s[sl][3][jy][jz] = s23.y;
                sm.s[sl][4][jy][jz] = at(sc, 6 * n + 4);
            }
        }
    };
    load(x0 - 1);
    load(x0);
    for (int x = x0; x < x1; ++x) {
        load(x + 1);
        __syncthreads();
        if (active) {
            unsigned xo[3], yo[3], zo[3];
            node_offsets(L, x, y, z, xo, yo, zo);
            Node<R, SHIFT> nd;
            // scalars of the source x - c_i from the ring
            auto scal = [&](int i, unsigned) {
                const int sl = (x - CX[i] + 9) % 3, jy = ty + 1 - CY[i], jz = tz + 1 - CZ[i];
                if constexpr (SHIFT)
                    return sm.s[sl][jy][jz];
                else
                    return FsScal64{sm.s[sl][0][jy][jz], sm.s[sl][1][jy][jz], sm.s[sl][2][jy][jz],
                                    sm.s[sl][3][jy][jz], sm.s[sl][4][jy][jz]};
            };
            const unsigned n = xo[1] + yo[1] + zo[1];
            if constexpr (EXT) {
                // special lanes (special_lane) are left to k_moments_special
                if (!special_lane(E, L, x, y, z, n)) {
                    const unsigned bb = wall_bb_mask(E, L, x, y, z);
                    const auto own = rev_scal(scal(0, 0u));
                    gather_fs<S, R, SHIFT>(
                        fs, N, xo, yo, zo, [&](int i, unsigned) { return ((bb >> i) & 1u) ? own : scal(i, 0u); }, nd,
                        R(L.rho0));
                    at(phi, n) = phi_entry<P>(moments<R, SHIFT>(L, nd));
                }
            } else {
                gather_fs<S, R, SHIFT>(fs, N, xo, yo, zo, scal, nd, R(L.rho0));
                const Moments<R, SHIFT> mo = moments<R, SHIFT>(L, nd);
                if constexpr (ENC)
                    at(phi, n) = P(minority_q(mo.sr, mo.sb, fdiv(R(1), mo.rho)));
                else
                    at(phi, n) = P(mo.phi);
            }
        }
        __syncthreads(); // the next load overwrites the oldest slot
    }
}

template <class S, class R, class P, bool SHIFT, bool ENC = false>
__global__ __launch_bounds__(256) void k_moments_fsx(Lat L, const void *__restrict__ p0, const void *__restrict__ p1,
                                                     const P *__restrict__ phio, P *__restrict__ phi, int xc)
{
    unsigned bx, by, bz;
    tile_of_block(L, bx, by, bz);
    moments_fsx_tile<S, R, P, SHIFT, ENC>(L, p0, p1, phio, phi, xc, bx, by, bz);
}

// K1 of the factorised layout with walls, mirrors and particles (all but the
// special lanes).
template <class S, class R, class P, bool SHIFT>
__global__ __launch_bounds__(256, 4) void k_moments_fsx_ext(Lat L, const void *__restrict__ p0,
                                                         const void *__restrict__ p1, const P *__restrict__ phio,
                                                         P *__restrict__ phi, int xc, Ext E)
{
    unsigned bx, by, bz;
    tile_of_block(L, bx, by, bz);
    moments_fsx_tile<S, R, P, SHIFT, false, true>(L, p0, p1, phio, phi, xc, bx, by, bz, E);
}

// ----------------------------------------------------- gradient of phi (b)
// Gradient 3*sum_i w_i c_i phi(x + c_i) over i = 1..18 in the reference's
// order (S:511-520), with zero terms skipped (they add exact zeros).
// The same stencil 3 sum_i w_i c_i v_i of the 18 neighbour values v[1..18]
// summed another way, for FP32: the values with their signs per weight class
// (axis and diagonal directions) first, the two weights once at the end.
// 30 operations instead of 51 (a multiply per term and an add or two); the
// sum is the same, rounded differently.  FP64 keeps the reference's order.
template <class R> __device__ __forceinline__ void stencil19(const R v[Q], R s[3])
{
    constexpr int PX = find_dir(1, 0, 0), MX = find_dir(-1, 0, 0), PY = find_dir(0, 1, 0), MY = find_dir(0, -1, 0);
    constexpr int PZ = find_dir(0, 0, 1), MZ = find_dir(0, 0, -1);
    constexpr int PPO = find_dir(1, 1, 0), MPO = find_dir(-1, 1, 0), PMO = find_dir(1, -1, 0), MMO = find_dir(-1, -1, 0);
    constexpr int POP = find_dir(1, 0, 1), MOP = find_dir(-1, 0, 1), POM = find_dir(1, 0, -1), MOM = find_dir(-1, 0, -1);
    constexpr int OPP = find_dir(0, 1, 1), OMP = find_dir(0, -1, 1), OPM = find_dir(0, 1, -1), OMM = find_dir(0, -1, -1);
    const R ax = v[PX] - v[MX], ay = v[PY] - v[MY], az = v[PZ] - v[MZ];
    const R bx = ((v[PPO] + v[PMO]) + (v[POP] + v[POM])) - ((v[MPO] + v[MMO]) + (v[MOP] + v[MOM]));
    const R by = ((v[PPO] + v[MPO]) + (v[OPP] + v[OPM])) - ((v[PMO] + v[MMO]) + (v[OMP] + v[OMM]));
    const R bz = ((v[POP] + v[MOP]) + (v[OPP] + v[OMP])) - ((v[POM] + v[MOM]) + (v[OPM] + v[OMM]));
    const R w1 = R(3.0 * W[PX]), w7 = R(3.0 * W[PPO]);
    s[0] = w1 * ax + w7 * bx;
    s[1] = w1 * ay + w7 * by;
    s[2] = w1 * az + w7 * bz;
}

template <class R, class F>
__device__ __forceinline__ void grad19(F phi_at, R &gx, R &gy, R &gz)
{
    if constexpr (sizeof(R) == 4) {
        R v[Q], s[3];
        v[0] = 0;
#pragma unroll
        for (int i = 1; i < Q; ++i) v[i] = R(phi_at(CX[i], CY[i], CZ[i]));
        stencil19<R>(v, s);
        gx = s[0];
        gy = s[1];
        gz = s[2];
        return;
    }
    gx = 0;
    gy = 0;
    gz = 0;
#pragma unroll
    for (int i = 1; i < Q; ++i) {
        R val = R(W[i]) * R(phi_at(CX[i], CY[i], CZ[i]));
        if (CX[i] == 1) gx = gx + val;
        if (CX[i] == -1) gx = gx - val;
        if (CY[i] == 1) gy = gy + val;
        if (CY[i] == -1) gy = gy - val;
        if (CZ[i] == 1) gz = gz + val;
        if (CZ[i] == -1) gz = gz - val;
    }
    gx = R(3) * gx;
    gy = R(3) * gy;
    gz = R(3) * gz;
}

// Normal derivatives for the curvature, d[k][l] = 3 sum_i w_i c_il
// n_k(x + c_i) (S:635-660), from the normals n_at(k, dx, dy, dz) around the
// node: FP32 through stencil19, FP64 in the reference's order.
template <class R, class NAt> __device__ __forceinline__ void curv19(NAt n_at, R d[3][3])
{
    if constexpr (sizeof(R) == 4) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            R v[Q];
            v[0] = 0;
#pragma unroll
            for (int i = 1; i < Q; ++i) v[i] = n_at(k, CX[i], CY[i], CZ[i]);
            stencil19<R>(v, d[k]);
        }
        return;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) d[k][0] = d[k][1] = d[k][2] = 0;
#pragma unroll
    for (int i = 1; i < Q; ++i) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            R tmp = R(W[i]) * n_at(k, CX[i], CY[i], CZ[i]);
            if (CX[i] == 1) d[k][0] = d[k][0] + tmp;
            if (CX[i] == -1) d[k][0] = d[k][0] - tmp;
            if (CY[i] == 1) d[k][1] = d[k][1] + tmp;
            if (CY[i] == -1) d[k][1] = d[k][1] - tmp;
            if (CZ[i] == 1) d[k][2] = d[k][2] + tmp;
            if (CZ[i] == -1) d[k][2] = d[k][2] - tmp;
        }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int l = 0; l < 3; ++l) d[k][l] = R(3) * d[k][l];
}

// Unit normal n = grad phi / |grad phi| and the interface flag |grad phi| >
// 1e-8 (S:498-532); n = 0 without the flag.  FP64 divides by the square root
// as the reference does, FP32 scales by the hardware reciprocal square root.
template <class R> __device__ __forceinline__ bool unit_normal(R gx, R gy, R gz, R &nx, R &ny, R &nz)
{
    const R g2 = gx * gx + gy * gy + gz * gz;
    nx = ny = nz = 0;
    if constexpr (sizeof(R) == 8) {
        const R fnorm = fsqrt(g2);
        if (!(fnorm > R(EVAP_LIM))) return false;
        nx = gx / fnorm;
        ny = gy / fnorm;
        nz = gz / fnorm;
    } else {
        if (!(g2 > R(EVAP_LIM * EVAP_LIM))) return false;
        const R inv = __builtin_amdgcn_rsqf(g2);
        nx = gx * inv;
        ny = gy * inv;
        nz = gz * inv;
    }
    return true;
}

// ------------------------------------------------------------- collision (a)
// CSF force from the curvature of the normals (S:624-667), velocity shift,
// viscosity interpolation and MRT relaxation with Guo forcing (S:669-689).
// In: the node's total populations f (FP64) or deviations f - w rho0
// (shifted), moments, phi, grad phi, n, the 3x3 normal derivatives.  Out: f
// replaced by the post-collision populations (or their deviations).
template <class R, bool SHIFT>
__device__ __forceinline__ void collide(const Lat &L, R f[Q], const Moments<R, SHIFT> &mo, R phin, R gx, R gy,
                                        R gz, R nx, R ny, R nz, bool force_on, const R d[3][3])
{
    R fx = 0, fy = 0, fz = 0;
    if (force_on) {
        R kurv = nx * ny * (d[1][0] + d[0][1]) + nx * nz * (d[2][0] + d[0][2]) + ny * nz * (d[1][2] + d[2][1]) -
                 (ny * ny + nz * nz) * d[0][0] - (nx * nx + nz * nz) * d[1][1] - (nx * nx + ny * ny) * d[2][2];
        R val = R(0.5) * R(L.sigma) * kurv;
        fx = val * gx;
        fy = val * gy;
        fz = val * gz;
    }
    const R rho = mo.rho;
    const R ux = mo.ux + fdiv(R(0.5) * fx, rho);
    const R uy = mo.uy + fdiv(R(0.5) * fy, rho);
    const R uz = mo.uz + fdiv(R(0.5) * fz, rho);
    R nu;
    if (phin <= R(-1))
        nu = R(L.nu_b);
    else if (phin >= R(1))
        nu = R(L.nu_r);
    else
        nu = fdiv(R(2), fdiv(R(1) + phin, R(L.nu_r)) + fdiv(R(1) - phin, R(L.nu_b)));
    const R su = fdiv(R(1), R(3) * nu + R(0.5));
    const R sq = fdiv(R(8) * (R(2) - su), R(8) - su);
    if constexpr (SHIFT) {
        // The relaxation rates split by moment parity: every even moment that
        // is not conserved relaxes with su and every odd one with sq (S:682),
        // the equilibrium moments are M feq of the SRT equilibrium and the
        // Guo forcing moments are M G (S:405-447).  Conserved moments (rho,
        // j) leave the same result at any rate because u carries the half
        // force.  So the MRT update is the two-relaxation-time update in
        // population space, per direction pair (i, opp i):
        //   f_i += -su (f+_i - feq+_i) - sq (f-_i - feq-_i)
        //          + (1 - su/2) G+_i + (1 - sq/2) G-_i
        // with f+- the half sum and difference of the pair.  Here on shifted
        // values f~ = f - w rho0 (the rho0 part is its own equilibrium).
        const R drho = mo.drho; // sum of the deviations (the moments' own sum)
        const R usqr = ux * ux + uy * uy + uz * uz;
        const R uf = ux * fx + uy * fy + uz * fz;
        const R ep0 = drho - R(1.5) * rho * usqr; // w-less even equilibrium part
        const R ap = R(1) - R(0.5) * su, am = R(1) - R(0.5) * sq;
        const R gp0 = R(-3) * uf;
        const R hsu = R(0.5) * su, hsq = R(0.5) * sq;
        // rest population: f0 += -su (f0 - feq0) + (1 - su/2) G0
        f[0] = f[0] - su * (f[0] - R(W[0]) * ep0) + ap * (R(W[0]) * gp0);
        // per weight class (w = 1/18 axis pairs, 1/36 diagonal pairs), with
        // feq+ = w (ep0 + 4.5 rho cu^2), feq- = 3 w rho cu,
        // G+ = w (gp0 + 9 cu cf), G- = 3 w cf:
        //   A = su w ep0 + ap w gp0 + (su w 4.5 rho) cu^2 + (9 ap w) cu cf - su/2 sp
        //   B = (3 sq w rho) cu + (3 am w) cf - sq/2 dm
        R kE[2], kQ[2], kX[2], kM[2], kF[2];
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const R w = R(c ? W[7] : W[1]);
            kE[c] = w * (su * ep0 + ap * gp0);
            kQ[c] = (w * su) * (R(4.5) * rho);
            kX[c] = (w * ap) * R(9);
            kM[c] = (w * sq) * (R(3) * rho);
            kF[c] = (w * am) * R(3);
        }
#pragma unroll
        for (int q = 0; q < NPAIR; ++q) {
            const int i = PAIR[q], o = OPC[i], c = i <= 6 ? 0 : 1;
            const R cu = dot_c<R>(i, ux, uy, uz);
            const R cf = dot_c<R>(i, fx, fy, fz);
            const R sp = f[i] + f[o], dm = f[i] - f[o];
            R A = kE[c] + kQ[c] * (cu * cu);
            A = A + kX[c] * (cu * cf);
            A = A - hsu * sp;
            R B = kM[c] * cu + kF[c] * cf;
            B = B - hsq * dm;
            f[i] = f[i] + (A + B);
            f[o] = f[o] + (A - B);
        }
        return;
    }
    R m[Q];
    if constexpr (!SHIFT) {
        // dense order of the reference's MATMUL(M, f) (S:685), zero terms skipped
#pragma unroll
        for (int k = 0; k < Q; ++k) {
            R acc = 0;
#pragma unroll
            for (int j = 0; j < Q; ++j) {
                const int c = MT.m[k][j];
                if (c == 0) continue;
                if (c == 1)
                    acc = acc + f[j];
                else if (c == -1)
                    acc = acc - f[j];
                else
                    acc = acc + R(c) * f[j];
            }
            m[k] = acc;
        }
    } else {
        // even moments from pair sums, odd moments from pair differences
        R ps[NPAIR], pd[NPAIR];
#pragma unroll
        for (int q = 0; q < NPAIR; ++q) {
            ps[q] = f[PAIR[q]] + f[OPC[PAIR[q]]];
            pd[q] = f[PAIR[q]] - f[OPC[PAIR[q]]];
        }
#pragma unroll
        for (int k = 0; k < Q; ++k) {
            const bool ev = MT.even[k];
            R acc = ev ? R(MT.m[k][0]) * f[0] : R(0);
#pragma unroll
            for (int q = 0; q < NPAIR; ++q) {
                const int c = MT.m[k][PAIR[q]];
                const R v = ev ? ps[q] : pd[q];
                if (c == 0) continue;
                if (c == 1)
                    acc += v;
                else if (c == -1)
                    acc -= v;
                else
                    acc += R(c) * v;
            }
            m[k] = acc;
        }
    }
    // equilibrium and forcing moments (equilf, S:405-447)
    R meq[Q], fk[Q];
    {
        R usqr = ux * ux + uy * uy + uz * uz;
        meq[0] = 1;
        meq[1] = R(19) * usqr - R(11);
        meq[2] = R(3) - R(5.5) * usqr;
        meq[3] = ux;
        meq[4] = div_c<R>(-ux, 1.5);
        meq[5] = uy;
        meq[6] = div_c<R>(-uy, 1.5);
        meq[7] = uz;
        meq[8] = div_c<R>(-uz, 1.5);
        meq[9] = R(3) * ux * ux - usqr;
        meq[10] = R(-0.5) * (R(3) * ux * ux - usqr);
        meq[11] = uy * uy - uz * uz;
        meq[12] = R(-0.5) * (uy * uy - uz * uz);
        meq[13] = ux * uy;
        meq[14] = uy * uz;
        meq[15] = ux * uz;
        meq[16] = 0;
        meq[17] = 0;
        meq[18] = 0;
        if constexpr (!SHIFT) {
#pragma unroll
            for (int k = 0; k < Q; ++k) meq[k] = rho * meq[k];
        } else {
            // deviation moments: meq - rho0 * (1, -11, 3, 0, ...)
            const R drho = m[0]; // the relaxed density stays exactly sum_i (f_i - w_i rho0)
            meq[0] = drho;
            meq[1] = R(19) * (rho * usqr) - R(11) * drho;
            meq[2] = R(3) * drho - R(5.5) * (rho * usqr);
#pragma unroll
            for (int k = 3; k < Q; ++k) meq[k] = rho * meq[k];
        }
        R uf = ux * fx + uy * fy + uz * fz;
        fk[0] = 0;
        fk[1] = R(38) * uf;
        fk[2] = R(-11) * uf;
        fk[3] = fx;
        fk[4] = div_c<R>(-fx, 1.5);
        fk[5] = fy;
        fk[6] = div_c<R>(-fy, 1.5);
        fk[7] = fz;
        fk[8] = div_c<R>(-fz, 1.5);
        fk[9] = R(6) * ux * fx - R(2) * uf;
        fk[10] = -(R(3) * ux * fx - uf);
        fk[11] = R(2) * (uy * fy - uz * fz);
        fk[12] = -(uy * fy - uz * fz);
        fk[13] = ux * fy + uy * fx;
        fk[14] = uy * fz + uz * fy;
        fk[15] = ux * fz + uz * fx;
        fk[16] = 0;
        fk[17] = 0;
        fk[18] = 0;
    }
    const R si[Q] = {R(1), su, su, R(1), sq, R(1), sq, R(1), sq, su, su, su, su, su, su, su, sq, sq, sq};
#pragma unroll
    for (int k = 0; k < Q; ++k) m[k] = m[k] - si[k] * (m[k] - meq[k]) + (R(1) - R(0.5) * si[k]) * fk[k];
    // back to populations: f_i = sum_k Minv[i][k] m_k
    if constexpr (!SHIFT) {
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            R acc = 0;
#pragma unroll
            for (int k = 0; k < Q; ++k) {
                if (MT.m[k][i] == 0) continue;
                acc = acc + R(MT.inv[i][k]) * m[k];
            }
            f[i] = acc;
        }
    } else {
        // scaled moments, then even and odd parts per direction pair
        R mh[Q];
#pragma unroll
        for (int k = 0; k < Q; ++k) mh[k] = m[k] * R(1.0 / MT.norm2[k]);
        {
            R acc = 0;
#pragma unroll
            for (int k = 0; k < Q; ++k)
                if (MT.m[k][0] != 0) acc += R(MT.m[k][0]) * mh[k];
            f[0] = acc;
        }
#pragma unroll
        for (int q = 0; q < NPAIR; ++q) {
            const int p = PAIR[q];
            R e = 0, o = 0;
#pragma unroll
            for (int k = 0; k < Q; ++k) {
                const int c = MT.m[k][p];
                if (c == 0) continue;
                R &acc = MT.even[k] ? e : o;
                if (c == 1)
                    acc += mh[k];
                else if (c == -1)
                    acc -= mh[k];
                else
                    acc += R(c) * mh[k];
            }
            f[p] = e + o;
            f[OPC[p]] = e - o;
        }
    }
}

// Recolouring parameters of a node (S:713-725): colour fraction A, the
// perturbation amplitude gamma (0 where the node is not flagged) and the
// EVAP_LIM overrides (r = 0, then r = f).
template <class R> struct Recol {
    R A, gam;
    bool evr, evb;
    R M;         // shifted storage: the minority colour's fraction (B where red is the majority, else A)
    bool redmaj; // shifted storage: rho_R >= rho_B
};
template <class R, bool SHIFT>
__device__ __forceinline__ Recol<R> recol(const Lat &L, const Moments<R, SHIFT> &mo, bool flag)
{
    Recol<R> c;
    if constexpr (SHIFT) {
        c.redmaj = mo.sb <= mo.sr;
        c.M = fdiv(c.redmaj ? mo.sb : mo.sr, mo.rho);
        c.A = c.redmaj ? R(1) - c.M : c.M;
    } else {
        c.A = fdiv(mo.sr, mo.rho);
    }
    c.gam = flag ? fdiv(mo.sr * mo.sb, mo.rho) * R(L.beta) : R(0);
    c.evr = mo.sr <= R(EVAP_LIM);
    c.evb = mo.sb <= R(EVAP_LIM);
    return c;
}

// Signed colour fraction (see decode_frac) after the EVAP_LIM overrides.
template <class R> __device__ __forceinline__ float encode_frac(const Recol<R> &c)
{
    if (c.evb) return -0.0f;
    if (c.evr) return 0.0f;
    return c.redmaj ? -float(c.M) : float(c.M);
}

// Output writers.  RB: the recoloured red and blue populations (S:713-726);
// shifted RB stores deviations from w_i rho0 s_red and w_i rho0 s_blue of the
// node that reads direction i next (its phase field phw_i), the node's
// minority colour formed directly and the majority colour as the complement
// of the total.  FS: the post-collision f and the node's recolour scalars.
template <int LOUT, class S, class R, bool SHIFT>
__device__ __forceinline__ void write_node(const Lat &L, void *o0, void *o1, unsigned n, const R f[Q],
                                           const Recol<R> &c, R nx, R ny, R nz, const R phw[Q])
{
    const size_t N = (size_t)L.N;
    if constexpr (LOUT == RB) {
        S *ro = (S *)o0, *bo = (S *)o1;
        if constexpr (!SHIFT) {
#pragma unroll
            for (int i = 0; i < Q; ++i) {
                R r = c.A * f[i];
                if (i > 0 && c.gam != R(0)) r = r + c.gam * R(W[i]) * dot_c<R>(i, nx, ny, nz);
                if (c.evr) r = 0;
                if (c.evb) r = f[i];
                const unsigned k = grp_off(i, N, n);
                StoreOf<S>::st(grp(ro, i, N), k, r);
                StoreOf<S>::st(grp(bo, i, N), k, f[i] - r);
            }
        } else {
            const R rho0 = R(L.rho0);
            const Shift<R> me{c.M, c.redmaj};
#pragma unroll
            for (int i = 0; i < Q; ++i) {
                const R dot = R(CX[i]) * nx + R(CY[i]) * ny + R(CZ[i]) * nz;
                const Shift<R> rd = shift_of<R>(phw[i]);
                const R wr = R(W[i]) * rho0;
                // r~ = A f~ + w rho0 (A - s_red) + g w n.c and b~ = f~ - r~, the
                // minority one formed first
                const R dsh = red_shift_diff(me, rd);
                const R gd = c.gam * R(W[i]) * dot;
                R vr, vb;
                if (c.redmaj) {
                    vb = c.M * f[i] - wr * dsh - gd;
                    vr = f[i] - vb;
                } else {
                    vr = c.M * f[i] + wr * dsh + gd;
                    vb = f[i] - vr;
                }
                const R sred = rd.redmaj ? R(1) - rd.hm : rd.hm, sblue = rd.redmaj ? rd.hm : R(1) - rd.hm;
                if (c.evr) { // r = 0
                    vr = -wr * sred;
                    vb = f[i] - vr;
                }
                if (c.evb) { // b = 0
                    vb = -wr * sblue;
                    vr = f[i] - vb;
                }
                const unsigned k = grp_off(i, N, n);
                StoreOf<S>::st(grp(ro, i, N), k, vr);
                StoreOf<S>::st(grp(bo, i, N), k, vb);
            }
        }
    } else {
        // the next state is read by the next step's kernels only: stream it
        // past the caches (FP32 values; FP16 and FP64 use plain stores)
        S *fo = (S *)o0;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            S *fp = grp(fo, i, N);
            const unsigned k = grp_off(i, N, n);
            if constexpr (SHIFT && sizeof(S) == 4)
                __builtin_nontemporal_store(float(f[i]), (float *)&at(fp, k));
            else
                StoreOf<S>::st(fp, k, f[i]);
        }
        R A = c.A, g = c.gam;
        if (c.evr) { A = 0; g = 0; }
        if (c.evb) { A = 1; g = 0; }
        if constexpr (!SHIFT) {
            double2 *sc = (double2 *)o1;
            at(sc, 3 * n) = make_double2(A, g);
            at(sc, 3 * n + 1) = make_double2(nx, ny);
            at(sc, 3 * n + 2) = make_double2(nz, 0.0);
        } else {
            typedef float f4v __attribute__((ext_vector_type(4)));
            const f4v v4 = {encode_frac(c), float(g * nx), float(g * ny), float(g * nz)};
            __builtin_nontemporal_store(v4, (f4v *)&at((float4 *)o1, n));
        }
    }
}

// ----------------------------------------------------------- K2 fused step
// One workgroup owns a TZ x TY column of (z, y) and marches x over
// [x0, x0 + xc).  Rings in LDS:
//   sphi[5]  phi planes x-2 .. x+2 with a 2-node halo in y and z
//   sn[3]    unit normals of planes x-1 .. x+1 with a 1-node halo (the
//            curvature stencil needs n at the 18 neighbours)
// Streaming is the pull gather from x - c_i (periodic wrap).
template <class R, class P, int KY = KTY<R>> struct K2Smem {
    P phi[5][KY + 4][TZ + 4];
    R n[3][3][KY + 2][TZ + 2];
};

__device__ __forceinline__ int slot5(int p) { return (p + 10) % 5; }
__device__ __forceinline__ int slot3(int p) { return (p + 9) % 3; }

// Akai wetting correction of a unit normal n at a live fluid node whose
// particle centre is c (S:541-582): the candidate closer to n of the two
// normals at contact angle cita to the wall normal nw.  FP64.
__device__ __forceinline__ void wetting(const Lat &L, const Ext &E, double xn, double yn, double zn, const double *c,
                                        double cita, double n[3])
{
    double nw[3] = {xn - c[0], yn - c[1], zn - c[2]};
    // minimum image on periodic axes (a no-op away from the seam)
    const int n3[3] = {L.nx, L.ny, L.nz};
#pragma unroll
    for (int a = 0; a < 3; ++a)
        if (!(E.wall[a][0] || E.wall[a][1] || E.mirror[a][0] || E.mirror[a][1])) {
            if (nw[a] > 0.5 * n3[a]) nw[a] -= n3[a];
            if (nw[a] < -0.5 * n3[a]) nw[a] += n3[a];
        }
    double st = sqrt(nw[0] * nw[0] + nw[1] * nw[1] + nw[2] * nw[2]);
    nw[0] /= st;
    nw[1] /= st;
    nw[2] /= st;
    double dmax = fabs(nw[0] - n[0]);
    if (fabs(nw[1] - n[1]) > dmax) dmax = fabs(nw[1] - n[1]);
    if (fabs(nw[2] - n[2]) > dmax) dmax = fabs(nw[2] - n[2]);
    const double theta = dmax < 1.0e-5 ? 1.0e-5 : acos(nw[0] * n[0] + nw[1] * n[1] + nw[2] * n[2]);
    const double ca = cos(cita), sa = sin(cita);
    const double a1 = ca - sa * cos(theta) / sin(theta), a2 = ca + sa * cos(theta) / sin(theta);
    const double b1 = sa / sin(theta), b2 = -sa / sin(theta);
    double v1[3], v2[3];
    for (int k = 0; k < 3; ++k) {
        v1[k] = a1 * nw[k] + b1 * n[k];
        v2[k] = a2 * nw[k] + b2 * n[k];
    }
    const double db1 = (v1[0] - n[0]) * (v1[0] - n[0]) + (v1[1] - n[1]) * (v1[1] - n[1]) + (v1[2] - n[2]) * (v1[2] - n[2]);
    const double db2 = (v2[0] - n[0]) * (v2[0] - n[0]) + (v2[1] - n[1]) * (v2[1] - n[1]) + (v2[2] - n[2]) * (v2[2] - n[2]);
    for (int k = 0; k < 3; ++k) n[k] = db1 < db2 ? v1[k] : v2[k];
}

// No curvature force next to a wall ghost layer (the guard z >= 3, x >= 3 of
// S:624, whose normals the gradient pass never writes).
__device__ __forceinline__ bool away_from_walls(const Ext &E, const Lat &L, int x, int y, int z)
{
    return !((E.wall[0][0] && x < 2) || (E.wall[0][1] && x > L.nx - 3) || (E.wall[1][0] && y < 2) ||
             (E.wall[1][1] && y > L.ny - 3) || (E.wall[2][0] && z < 2) || (E.wall[2][1] && z > L.nz - 3));
}

// The update of one live node (x, y, z) of the fused step (derivatives,
// collision, recolouring, streaming as the pull gather), given the phase
// field phi_at(dx, dy, dz) and the unit normals n_at(k, dx, dy, dz) around it
// and, for a factorised input, the recolour scalars scal_at(o) of the node at
// offset o: from the tile's LDS rings in k_collide, from memory in
// k_collide_special.  WALLS (factorised tile kernel with walls, mirrors or
// particles): wbb holds the node's wall bounce-back directions (rev_scal);
// such a node is next to a wall, where no curvature force acts (S:624).
template <int LIN, int LOUT, class S, class R, class P, bool SHIFT, bool EXT, bool WALLS = false, class PhiAt,
          class NAt, class ScalAt>
__device__ __forceinline__ void lane_step(const Lat &L, const void *__restrict__ i0, const void *__restrict__ i1,
                                          void *__restrict__ o0, void *__restrict__ o1, const P *__restrict__ phio,
                                          const Ext &E, int x, int y, int z, PhiAt phi_at, NAt n_at, ScalAt scal_at,
                                          unsigned wbb = 0)
{
    unsigned nt = 0, bb = 0, madj = 0;
    if constexpr (EXT) {
        if (is_ghost(E, L, x, y, z)) return; // ghost layers are not collided
        bb = wall_bb_mask(E, L, x, y, z);
        madj = mirror_adjacent(E, L, x, y, z);
    }

    // gradient, normal and flag at the node (S:498-532)
    R gx, gy, gz;
    grad19<R>([&](int dx, int dy, int dz) { return phi_at(dx, dy, dz); }, gx, gy,
              gz);
    const R g2 = gx * gx + gy * gy + gz * gz;
    bool flag = sizeof(R) == 8 ? fsqrt(g2) > R(EVAP_LIM) : g2 > R(EVAP_LIM * EVAP_LIM);
    const R nx = n_at(0, 0, 0, 0);
    const R ny = n_at(1, 0, 0, 0);
    const R nz = n_at(2, 0, 0, 0);

    unsigned xo[3], yo[3], zo[3];
    node_offsets(L, x, y, z, xo, yo, zo);
    const unsigned n = xo[1] + yo[1] + zo[1];
    bool force_on = flag && !wbb;
    if constexpr (EXT) {
        if (E.ntype) nt = at(E.ntype, n);
        if (nt & NT_SOLID) flag = false;
        if (flag && (nt & NT_POUT)) { // wetting rescales the gradient too (S:569-581)
            const R fnorm = fsqrt(g2);
            gx = fnorm * nx;
            gy = fnorm * ny;
            gz = fnorm * nz;
        }
        force_on = flag && !(nt & NT_POUT) && away_from_walls(E, L, x, y, z);
        if (nt & NT_NEAR) {
            const size_t N = (size_t)L.N;
            const bool fl = flag;
            E.der[n] = fl ? double(gx) : 0.0;
            E.der[N + n] = fl ? double(gy) : 0.0;
            E.der[2 * N + n] = fl ? double(gz) : 0.0;
            E.der[3 * N + n] = double(nx);
            E.der[4 * N + n] = double(ny);
            E.der[5 * N + n] = double(nz);
            if (!(nt & NT_SOLID) && !fl) { // non-interface fluid: gradient kept, normal zero (S:519-530)
                E.der[n] = double(gx);
                E.der[N + n] = double(gy);
                E.der[2 * N + n] = double(gz);
            }
        }
    }
    // normal derivatives for the curvature (S:635-660), from the normals
    // ring; with walls and particles formed after the gather, so that they
    // are not live across it
    R d[3][3];
    auto curv_d = [&]() {
        if (force_on) curv19<R>(n_at, d);
    };
    if constexpr (!EXT) curv_d();

    R ax = 0, aw[Q]; // phase fields fixing the shift fractions: this node's of the stored
                     // state, and aw_i of the node that reads direction i next
    if constexpr (SHIFT && LIN == RB) ax = R(at(phio, n));
    if constexpr (SHIFT && LOUT == RB) {
        // shift of the slot read by x + c_i at the next step: that node's phi
        // now; a slot that bounces back off a wall is read by x itself
        const R aself = R(phi_at(0, 0, 0));
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            aw[i] = R(phi_at(CX[i], CY[i], CZ[i]));
            if constexpr (EXT)
                if ((bb >> OPC[i]) & 1u) aw[i] = aself;
        }
    }
    if constexpr (EXT && LIN == RB && LOUT == RB) {
        if (nt & NT_SOLID) {
            // solid nodes are not collided: their streamed populations pass
            // through to the next streaming (S:622, S:712, S:743)
            // (pulled as the live nodes pull: bounce-back off walls, the
            // mirror image across mirror faces)
            const size_t N = (size_t)L.N;
            const S *rs = (const S *)i0, *bs = (const S *)i1;
            S *ro = (S *)o0, *bo = (S *)o1;
            static_for<Q>([&](auto ic) {
                constexpr int i = decltype(ic)::value;
                unsigned m;
                int dsrc;
                pull_slot<i>(xo, yo, zo, bb, madj, m, dsrc);
                const unsigned ks = grp_off(dsrc, N, m), kd = grp_off(i, N, n);
                R rv = StoreOf<S>::ld(grp(rs, dsrc, N), ks), bv = StoreOf<S>::ld(grp(bs, dsrc, N), ks);
                if constexpr (SHIFT) {
                    const R dsh = R(W[i]) * R(L.rho0) * red_shift_diff(shift_of<R>(ax), shift_of<R>(aw[i]));
                    rv += dsh;
                    bv -= dsh;
                }
                StoreOf<S>::st(grp(ro, i, N), kd, rv);
                StoreOf<S>::st(grp(bo, i, N), kd, bv);
            });
            return;
        }
    }
    if constexpr (EXT && !(LIN == RB && LOUT == RB)) {
        if (nt & NT_SOLID) {
            // solid nodes are not collided (S:622, S:712, S:743): each slot
            // takes the total it pulls, and that slot's colours go to the
            // side arrays of the output; FS_SOLID scalars mark the node
            const size_t N = (size_t)L.N;
            const R rho0 = R(L.rho0);
            S *fo = (S *)o0;
            static_for<Q>([&](auto ic) {
                constexpr int i = decltype(ic)::value;
                const unsigned kd = grp_off(i, N, n);
                R f, vr = 0, vb = 0;
                float q = 0.f;
                if constexpr (LIN == RB) {
                    // the red/blue state of the first step, pulled as the
                    // live nodes pull it (bounce-back, mirror images)
                    unsigned m;
                    int d;
                    pull_slot<i>(xo, yo, zo, bb, madj, m, d);
                    const unsigned ks = grp_off(d, N, m);
                    const R rv = StoreOf<S>::ld(grp((const S *)i0, d, N), ks);
                    const R bv = StoreOf<S>::ld(grp((const S *)i1, d, N), ks);
                    f = rv + bv;
                    if constexpr (SHIFT) {
                        // deviations from this node's shifts: the absolute
                        // colours, the minority's from its own small terms
                        const Shift<R> sh = shift_of<R>(R(at(phio, n)));
                        const R wr = R(W[i]) * rho0, sred = sh.redmaj ? R(1) - sh.hm : sh.hm;
                        const R sblue = sh.redmaj ? sh.hm : R(1) - sh.hm;
                        q = enc_slot(rv + wr * sred, bv + wr * sblue);
                    } else {
                        vr = rv;
                        vb = bv;
                    }
                } else {
                    const unsigned m = xo[1 - CX[i]] + yo[1 - CY[i]] + zo[1 - CZ[i]], ks = grp_off(i, N, m);
                    int o[3];
                    unsigned flip;
                    fs_src<i>(bb, madj, o, flip);
                    f = StoreOf<S>::ld(grp((const S *)i0, i, N), ks);
                    slot_side<i, S, R, SHIFT>(E.rbi[0], E.rbi[1], N, ks, f, scal_at(o), flip, rho0, vr, vb, q);
                }
                StoreOf<S>::st(grp(fo, i, N), kd, f);
                if constexpr (SHIFT) {
                    at(grp((float *)E.rbo[0], i, N), kd) = q;
                } else {
                    StoreOf<S>::st(grp((S *)E.rbo[0], i, N), kd, vr);
                    StoreOf<S>::st(grp((S *)E.rbo[1], i, N), kd, vb);
                }
            });
            if constexpr (SHIFT) {
                at((float4 *)o1, n) = make_float4(FS_SOLID, 0.f, 0.f, 0.f);
            } else {
                double2 *sc = (double2 *)o1;
                at(sc, 3 * n) = make_double2(double(FS_SOLID), 0.0);
                at(sc, 3 * n + 1) = make_double2(0.0, 0.0);
                at(sc, 3 * n + 2) = make_double2(0.0, 0.0);
            }
            return;
        }
    }
    Node<R, SHIFT> nd;
    Moments<R, SHIFT> mo;
    bool refilled = false;
    if constexpr (EXT) refilled = (nt & NT_REFILL) != 0;
    if (refilled) {
        // refilled node (S:1364-1391): SRT equilibria of the refill densities
        // and the rigid-body velocity; macros as refilling set them
        const size_t N = (size_t)L.N;
        if constexpr (!SHIFT) {
            const double rr = E.mac[N + n], rbv = E.mac[2 * N + n];
            const double ux = E.mac[3 * N + n], uy = E.mac[4 * N + n], uz = E.mac[5 * N + n];
            const double usqr = ux * ux + uy * uy + uz * uz;
#pragma unroll
            for (int i = 0; i < Q; ++i) {
                const double uc = ux * CX[i] + uy * CY[i] + uz * CZ[i];
                const double poly = (1.0 + 3.0 * uc + 4.5 * uc * uc - 1.5 * usqr);
                nd.r[i] = rr * W[i] * poly;
                nd.b[i] = rbv * W[i] * poly;
            }
            mo.ux = R(ux);
            mo.uy = R(uy);
            mo.uz = R(uz);
        } else {
            // shifted storage: the deviations w_i (rho poly_i - rho0), in
            // the working precision
            const R rt = R(E.mac[N + n] + E.mac[2 * N + n]);
            const R ux = R(E.mac[3 * N + n]), uy = R(E.mac[4 * N + n]), uz = R(E.mac[5 * N + n]);
            const R usqr = ux * ux + uy * uy + uz * uz;
            R s = 0;
#pragma unroll
            for (int i = 0; i < Q; ++i) {
                const R uc = ux * R(CX[i]) + uy * R(CY[i]) + uz * R(CZ[i]);
                const R poly = R(1) + R(3) * uc + R(4.5) * uc * uc - R(1.5) * usqr;
                nd.ft[i] = R(W[i]) * (rt * poly - R(L.rho0));
                s += nd.ft[i];
            }
            mo.drho = s;
            mo.ux = ux;
            mo.uy = uy;
            mo.uz = uz;
        }
        mo.sr = R(E.mac[N + n]);
        mo.sb = R(E.mac[2 * N + n]);
        mo.rho = R(E.mac[n]);
        mo.phi = R(phi_at(0, 0, 0));
    } else {
        const size_t N = (size_t)L.N;
        if constexpr (EXT && LIN == FS) {
            // slots split by the scalars of their node, of the node itself off
            // a wall or of a mirror image, or taken from the side arrays of
            // solid nodes
            gather_fs_ext<true, S, R, SHIFT>((const S *)i0, E.rbi[0], E.rbi[1], N, xo, yo, zo, bb, madj, scal_at, nd,
                                             R(L.rho0));
        } else if constexpr (LIN == FS && (SHIFT || WALLS)) {
            // the periodic gather; next to a wall (wbb) the bounce-back slots
            // are split with the node's own scalars (rev_scal)
            const int oc[3] = {0, 0, 0};
            const auto own = rev_scal(scal_at(oc));
            gather_fs<S, R, SHIFT>(
                (const S *)i0, N, xo, yo, zo,
                [&](int i, unsigned) {
                    const int o[3] = {-CX[i], -CY[i], -CZ[i]};
                    return ((wbb >> i) & 1u) ? own : scal_at(o);
                },
                nd, R(L.rho0));
        } else {
            gather<LIN, S, R, SHIFT>(L, i0, i1, xo, yo, zo, ax, nd, bb, madj);
        }
        mo = moments<R, SHIFT>(L, nd);
    }
    const R phin = R(phi_at(0, 0, 0));
    R f[Q];
    if constexpr (!SHIFT) {
#pragma unroll
        for (int i = 0; i < Q; ++i) f[i] = nd.r[i] + nd.b[i];
    } else {
#pragma unroll
        for (int i = 0; i < Q; ++i) f[i] = nd.ft[i];
    }
    if constexpr (EXT) curv_d();
    collide<R, SHIFT>(L, f, mo, phin, gx, gy, gz, nx, ny, nz, force_on, d);
    const Recol<R> rc = recol<R, SHIFT>(L, mo, flag);
    write_node<LOUT, S, R, SHIFT>(L, o0, o1, n, f, rc, nx, ny, nz, aw);
}

// The (z-tile bx, y-tile by, x-chunk bz) column of one workgroup.  SKIP
// (factorised layout with walls, mirrors or particles): the rings and the
// plain update of every lane but the special ones (special_lane), which
// k_collide_special updates.
template <int LIN, int LOUT, class S, class R, class P, bool SHIFT, bool EXT, bool SKIP = false>
__device__ __forceinline__ void collide_tile(const Lat &L, const void *__restrict__ i0, const void *__restrict__ i1,
                                             void *__restrict__ o0, void *__restrict__ o1,
                                             const P *__restrict__ phi, const P *__restrict__ phio, int xc,
                                             const Ext &E, unsigned bx, unsigned by, unsigned bz)
{
    __shared__ K2Smem<R, P> sm;
    const int tz = threadIdx.x, ty = threadIdx.y, tid = ty * TZ + tz;
    const int z0 = bx * TZ, y0 = by * KTY<R>, x0 = bz * xc;
    const int x1 = min(x0 + xc, L.nx);
    const int z = z0 + tz, y = y0 + ty;
    const bool active = z < L.nz && y < L.ny;
    const int sx = L.ny * L.nz;

    auto load_phi = [&](int p) {
        const int gx = wrapm(p, L.nx) * sx;
        P(*dst)[TZ + 4] = sm.phi[slot5(p)];
        for (int e = tid; e < (KTY<R> + 4) * (TZ + 4); e += TZ * KTY<R>) {
            int jy = e / (TZ + 4), jz = e - jy * (TZ + 4);
            int gy = wrapm(y0 - 2 + jy, L.ny), gz = wrapm(z0 - 2 + jz, L.nz);
            dst[jy][jz] = at(phi, (unsigned)(gx + gy * L.nz + gz));
        }
    };
    // unit normal of plane p on the (KTY<R>+2) x (TZ+2) halo region (S:498-532);
    // with EXT: zero at solid and ghost nodes, wetting at live fluid nodes
    // (planes plo..phi are in the phi ring)
    auto normals = [&](int p, int plo, int phi_) {
        const int s = slot3(p);
        for (int e = tid; e < (KTY<R> + 2) * (TZ + 2); e += TZ * KTY<R>) {
            int jy = e / (TZ + 2), jz = e - jy * (TZ + 2);
            R nx = 0, ny = 0, nz = 0;
            bool skip = false;
            int gxw = 0, gyw = 0, gzw = 0;
            unsigned nt = 0, A = 0;
            int ox = 0, oy = 0, oz = 0; // a mirror ghost takes its image's normal, reflected (S:589-591)
            if constexpr (EXT) {
                gxw = wrapm(p, L.nx);
                gyw = wrapm(y0 - 1 + jy, L.ny);
                gzw = wrapm(z0 - 1 + jz, L.nz);
                skip = is_wall_ghost(E, L, gxw, gyw, gzw);
                if (!skip) {
                    A = mirror_axes(E, L, gxw, gyw, gzw);
                    if (A) {
                        ox = (A & 1u) ? (gxw == 0 ? 1 : -1) : 0;
                        oy = (A & 2u) ? (gyw == 0 ? 1 : -1) : 0;
                        oz = (A & 4u) ? (gzw == 0 ? 1 : -1) : 0;
                        // the image's stencil must lie in the rings (else no live node uses it)
                        skip = p + ox - 1 < plo || p + ox + 1 > phi_ || jy + oy < 0 || jy + oy > KTY<R> + 1 ||
                               jz + oz < 0 || jz + oz > TZ + 1;
                    }
                }
                if (!skip && E.ntype) {
                    nt = at(E.ntype, (unsigned)(gxw * sx + gyw * L.nz + gzw));
                    skip = (nt & NT_SOLID) != 0;
                }
            }
            if (!skip) {
                R gx, gy, gz;
                grad19<R>(
                    [&](int dx, int dy, int dz) {
                        return sm.phi[slot5(p + ox + dx)][jy + 1 + oy + dy][jz + 1 + oz + dz];
                    },
                    gx, gy, gz);
                const bool fl = unit_normal<R>(gx, gy, gz, nx, ny, nz);
                if constexpr (EXT) {
                    if (fl && (nt & NT_POUT)) {
                        // the wetted normal that k_wet_normals stored for the
                        // node, or for a mirror ghost for its image (whose
                        // masks the ghost holds)
                        const size_t N = (size_t)L.N;
                        const unsigned im = (unsigned)((gxw + ox) * sx + (gyw + oy) * L.nz + gzw + oz);
                        nx = R(E.der[3 * N + im]);
                        ny = R(E.der[4 * N + im]);
                        nz = R(E.der[5 * N + im]);
                    }
                    if (A & 1u) nx = -nx;
                    if (A & 2u) ny = -ny;
                    if (A & 4u) nz = -nz;
                }
            }
            sm.n[0][s][jy][jz] = nx;
            sm.n[1][s][jy][jz] = ny;
            sm.n[2][s][jy][jz] = nz;
        }
    };

    // FP32/FP16 factorised input: the recolour scalars of planes x-1..x+1 in an
    // LDS ring, as in the x-marching moments kernel (the FP64 ring would not
    // leave room for a second workgroup per CU)
    constexpr bool RING = LIN == FS && SHIFT;
    struct NoRing {};
    __shared__ std::conditional_t<RING, FsxRing<true, KTY<R>>, NoRing> ring;
    auto load_scal = [&](int p) {
        if constexpr (RING) {
            const int sl = (p + 9) % 3;
            const unsigned gx = (unsigned)wrapm(p, L.nx) * (unsigned)sx;
            for (int e = tid; e < (KTY<R> + 2) * (TZ + 2); e += TZ * KTY<R>) {
                const int jy = e / (TZ + 2), jz = e - jy * (TZ + 2);
                const unsigned n =
                    gx + (unsigned)wrapm(y0 - 1 + jy, L.ny) * (unsigned)L.nz + wrapm(z0 - 1 + jz, L.nz);
                ring.s[sl][jy][jz] = at((const float4 *)i1, n);
            }
        }
    };

    for (int p = x0 - 2; p <= x0 + 2; ++p) load_phi(p);
    load_scal(x0 - 1);
    load_scal(x0);
    __syncthreads();
    normals(x0 - 1, x0 - 2, x0 + 2);
    normals(x0, x0 - 2, x0 + 2);

    for (int x = x0; x < x1; ++x) {
        __syncthreads(); // readers of the slots refilled below are done
        if (x > x0) load_phi(x + 2);
        load_scal(x + 1);
        __syncthreads();
        normals(x + 1, x - 2, x + 2);
        __syncthreads();
        if (!active) continue;
        unsigned wbb = 0;
        if constexpr (SKIP) {
            if (special_lane(E, L, x, y, z, ((unsigned)x * L.ny + y) * L.nz + z)) continue;
            wbb = wall_bb_mask(E, L, x, y, z);
        }
        lane_step<LIN, LOUT, S, R, P, SHIFT, EXT, SKIP>(
            L, i0, i1, o0, o1, phio, E, x, y, z,
            [&](int dx, int dy, int dz) { return sm.phi[slot5(x + dx)][ty + 2 + dy][tz + 2 + dz]; },
            [&](int k, int dx, int dy, int dz) { return sm.n[k][slot3(x + dx)][ty + 1 + dy][tz + 1 + dz]; },
            [&](const int o[3]) {
                if constexpr (RING)
                    return ring.s[(x + o[0] + 9) % 3][ty + 1 + o[1]][tz + 1 + o[2]];
                else if constexpr (LIN == FS && !SHIFT)
                    return fs64_at((const double *)i1,
                                   ((unsigned)wrapm(x + o[0], L.nx) * L.ny + wrapm(y + o[1], L.ny)) * L.nz +
                                       wrapm(z + o[2], L.nz));
                else
                    return 0;
            },
            wbb);
    }
}

template <int LIN, int LOUT, class S, class R, class P, bool SHIFT, bool EXT = false, bool SKIP = false>
__global__ __launch_bounds__(TZ * KTY<R>, KWAVES<R>) void k_collide(Lat L, const void *__restrict__ i0, const void *__restrict__ i1,
                                                 void *__restrict__ o0, void *__restrict__ o1,
                                                 const P *__restrict__ phi, const P *__restrict__ phio, int xc,
                                                 Ext E = Ext{})
{
    unsigned bx, by, bz;
    tile_of_block(L, bx, by, bz);
    collide_tile<LIN, LOUT, S, R, P, SHIFT, EXT, SKIP>(L, i0, i1, o0, o1, phi, phio, xc, E, bx, by, bz);
}

// ------------------------------------------------------------- diagnostics
// FP64 sums over the live fluid nodes of the macroscopic fields that the
// moments kernel wrote (rho, rho_R, rho_B, ux, uy, uz, phi; 7 x N): mass,
// red and blue mass, momentum and kinetic energy.  Each workgroup reduces a
// fixed range of nodes in a fixed tree order; k_reduce_final adds the
// partials in workgroup order, so the sums are reproducible.
constexpr int RED_B = 256;
constexpr int NDIAG = 7;

__device__ __forceinline__ bool live_fluid(const Ext &E, const Lat &L, unsigned n)
{
    const int z = (int)(n % (unsigned)L.nz), y = (int)((n / (unsigned)L.nz) % (unsigned)L.ny),
              x = (int)(n / ((unsigned)L.nz * (unsigned)L.ny));
    if (is_ghost(E, L, x, y, z)) return false;
    return !(E.ntype && (E.ntype[n] & NT_SOLID));
}

static __global__ __launch_bounds__(RED_B) void k_reduce(Lat L, Ext E, const double *__restrict__ mac,
                                                  double *__restrict__ part, unsigned per_block)
{
    __shared__ double sh[NDIAG][RED_B];
    const size_t N = (size_t)L.N;
    double acc[NDIAG] = {0, 0, 0, 0, 0, 0, 0};
    const unsigned n0 = blockIdx.x * per_block, n1 = min((unsigned)N, n0 + per_block);
    for (unsigned n = n0 + threadIdx.x; n < n1; n += RED_B) {
        if (!live_fluid(E, L, n)) continue;
        const double rho = mac[n], ux = mac[3 * N + n], uy = mac[4 * N + n], uz = mac[5 * N + n];
        acc[0] += rho;
        acc[1] += mac[N + n];
        acc[2] += mac[2 * N + n];
        acc[3] += rho * ux;
        acc[4] += rho * uy;
        acc[5] += rho * uz;
        acc[6] += 0.5 * rho * (ux * ux + uy * uy + uz * uz);
    }
    for (int k = 0; k < NDIAG; ++k) sh[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int st = RED_B / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st)
            for (int k = 0; k < NDIAG; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x < NDIAG) part[(size_t)blockIdx.x * NDIAG + threadIdx.x] = sh[threadIdx.x][0];
}

static __global__ __launch_bounds__(64) void k_reduce_final(const double *__restrict__ part, unsigned nblocks,
                                                     double *__restrict__ out)
{
    if (threadIdx.x >= NDIAG) return;
    double s = 0.0;
    for (unsigned b = 0; b < nblocks; ++b) s += part[(size_t)b * NDIAG + threadIdx.x];
    out[threadIdx.x] = s;
}

// Red area of each z slice: sum over the slice's live fluid nodes of
// (1 + phi) / 2, one workgroup per slice (the neck radius of a jet along z is
// sqrt(area / pi)).
static __global__ __launch_bounds__(RED_B) void k_slice_red(Lat L, Ext E, const double *__restrict__ mac,
                                                     double *__restrict__ area)
{
    __shared__ double sh[RED_B];
    const size_t N = (size_t)L.N;
    const unsigned z = blockIdx.x, nxy = (unsigned)L.nx * (unsigned)L.ny;
    double acc = 0.0;
    for (unsigned k = threadIdx.x; k < nxy; k += RED_B) {
        const unsigned n = k * (unsigned)L.nz + z; // (x, y) = (k / ny, k % ny)
        if (live_fluid(E, L, n)) acc += 0.5 * (1.0 + mac[6 * N + n]);
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int st = RED_B / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) area[z] = sh[0];
}

// Diagnostics without a field buffer: each thread forms the moments of its
// (y, z) column's nodes as calmacros does (node_moments) while marching x, and
// sums over the live fluid nodes the k_reduce quantities and the red area
// (1 + phi)/2 of its column.  A workgroup (one 64 x 4 tile of (z, y)) writes
// its NDIAG sums to part[tile] (fixed-order tree) and the red area of its
// 64 z slices summed over its 4 rows to spart[y tile][z]; k_diag_final adds
// them in a fixed order.  Nothing but the two partial arrays is written.
template <int LIN, class S, class R, class P, bool SHIFT, bool EXT>
__global__ __launch_bounds__(TZ *TY) void k_diag_tiles(Lat L, const void *__restrict__ p0,
                                                        const void *__restrict__ p1, const P *__restrict__ phio,
                                                        Ext E, double *__restrict__ part, double *__restrict__ spart)
{
    __shared__ double sh[NDIAG][TZ * TY];
    __shared__ double sa[TY][TZ];
    const int tz = threadIdx.x, ty = threadIdx.y, tid = ty * TZ + tz;
    const int z = blockIdx.x * TZ + tz, y = blockIdx.y * TY + ty;
    double acc[NDIAG] = {0, 0, 0, 0, 0, 0, 0}, area = 0.0;
    if (z < L.nz && y < L.ny) {
        for (int x = 0; x < L.nx; ++x) {
            unsigned n, nt;
            Moments<R, SHIFT> mo;
            if (!node_moments<LIN, S, R, P, SHIFT, EXT>(L, p0, p1, phio, E, x, y, z, n, mo, nt)) continue;
            if (nt & NT_SOLID) continue;
            const double rho = double(mo.rho), ux = double(mo.ux), uy = double(mo.uy), uz = double(mo.uz);
            acc[0] += rho;
            acc[1] += double(mo.sr);
            acc[2] += double(mo.sb);
            acc[3] += rho * ux;
            acc[4] += rho * uy;
            acc[5] += rho * uz;
            acc[6] += 0.5 * rho * (ux * ux + uy * uy + uz * uz);
            area += 0.5 * (1.0 + double(mo.phi));
        }
    }
    for (int k = 0; k < NDIAG; ++k) sh[k][tid] = acc[k];
    sa[ty][tz] = area;
    __syncthreads();
    for (int st = TZ * TY / 2; st > 0; st >>= 1) {
        if (tid < st)
            for (int k = 0; k < NDIAG; ++k) sh[k][tid] += sh[k][tid + st];
        __syncthreads();
    }
    const unsigned tile = blockIdx.y * gridDim.x + blockIdx.x;
    if (tid < NDIAG) part[(size_t)tile * NDIAG + tid] = sh[tid][0];
    if (ty == 0 && z < L.nz) {
        double s = 0.0;
        for (int r = 0; r < TY; ++r) s += sa[r][tz];
        spart[(size_t)blockIdx.y * L.nz + z] = s;
    }
}

// Fixed-order sums of k_diag_tiles' partials: the NDIAG totals into out and
// the red area of every z slice into area.
static __global__ __launch_bounds__(256) void k_diag_final(const double *__restrict__ part, unsigned ntiles,
                                                           const double *__restrict__ spart, unsigned gy, int nz,
                                                           double *__restrict__ out, double *__restrict__ area)
{
    if (threadIdx.x < NDIAG) {
        double s = 0.0;
        for (unsigned t = 0; t < ntiles; ++t) s += part[(size_t)t * NDIAG + threadIdx.x];
        out[threadIdx.x] = s;
    }
    for (int z = threadIdx.x; z < nz; z += blockDim.x) {
        double s = 0.0;
        for (unsigned b = 0; b < gy; ++b) s += spart[(size_t)b * nz + z];
        area[z] = s;
    }
}

// ------------------------------------------------- batches of independent boxes
// Boxes are packed one after another in the node index: box b owns nodes
// [base, base + N_b) of every per-node array and its own SoA block of each
// population array.  One launch covers the tiles of all boxes; a workgroup
// finds its box by binary search over the first-tile table.
struct BoxDev {
    Lat L;
    unsigned long long base; // first node of the box in the packed index
    unsigned t1, t2;         // first tile of the box in the moments / collide launch
    unsigned gz, gy, gx1;    // z tiles, y tiles, x chunks of the moments launch
    unsigned gy2, gx2;       // y tiles, x chunks of the collide launch
    unsigned pad;
};

__device__ __forceinline__ unsigned xcd_order(unsigned lin, unsigned nb)
{
    const unsigned full = nb - nb % NXCD;
    return lin < full ? (lin % NXCD) * (full / NXCD) + lin / NXCD : lin;
}

template <bool K2> __device__ __forceinline__ int find_box(const BoxDev *__restrict__ B, int nbox, unsigned t)
{
    int lo = 0, hi = nbox - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        const unsigned first = K2 ? B[mid].t2 : B[mid].t1;
        if (first <= t)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// Offset the two arrays of a state layout to box `base`.
template <int LAY, class S, bool SHIFT>
__device__ __forceinline__ void box_state(const void *&a, const void *&b, unsigned long long base)
{
    a = (const S *)a + (size_t)Q * base;
    if constexpr (LAY == RB)
        b = (const S *)b + (size_t)Q * base;
    else if constexpr (SHIFT)
        b = (const float4 *)b + base;
    else
        b = (const double *)b + (size_t)FsScal<false>::NS * base;
}

// Moments over x chunks of xc planes: the x-marching kernel for the
// factorised layout, plane by plane for red/blue.
template <int LIN, class S, class R, class P, bool SHIFT, bool ENC = false>
__global__ __launch_bounds__(256) void k_moments_batch(const BoxDev *__restrict__ B, int nbox, int remap, int xc,
                                                       const void *__restrict__ p0, const void *__restrict__ p1,
                                                       const P *__restrict__ phio, P *__restrict__ phi)
{
    const unsigned t = remap ? xcd_order(blockIdx.x, gridDim.x) : blockIdx.x;
    const int b = find_box<false>(B, nbox, t);
    const Lat L = B[b].L;
    const unsigned long long base = B[b].base;
    const unsigned lt = t - B[b].t1, gz = B[b].gz, gy = B[b].gy;
    const unsigned bx = lt % gz, by = (lt / gz) % gy, cx = lt / (gz * gy);
    const void *a = p0, *c = p1;
    box_state<LIN, S, SHIFT>(a, c, base);
    if constexpr (LIN == FS) {
        moments_fsx_tile<S, R, P, SHIFT, ENC>(L, a, c, phio + base, phi + base, xc, bx, by, cx);
    } else {
        const int x1 = min((int)(cx + 1) * xc, L.nx);
        for (int x = cx * xc; x < x1; ++x)
            moments_node<LIN, S, R, P, SHIFT, false, false>(L, a, c, phio + base, phi + base, nullptr, Ext{}, bx, by,
                                                            x);
    }
}

template <int LIN, int LOUT, class S, class R, class P, bool SHIFT>
__global__ __launch_bounds__(TZ * KTY<R>, KWAVES<R>) void k_collide_batch(const BoxDev *__restrict__ B, int nbox, int remap, int xc,
                                                       const void *__restrict__ i0, const void *__restrict__ i1,
                                                       void *__restrict__ o0, void *__restrict__ o1,
                                                       const P *__restrict__ phi, const P *__restrict__ phio)
{
    const unsigned t = remap ? xcd_order(blockIdx.x, gridDim.x) : blockIdx.x;
    const int b = find_box<true>(B, nbox, t);
    const Lat L = B[b].L;
    const unsigned long long base = B[b].base;
    const unsigned lt = t - B[b].t2, gz = B[b].gz, gy = B[b].gy2;
    const void *a0 = i0, *a1 = i1, *c0 = o0, *c1 = o1;
    box_state<LIN, S, SHIFT>(a0, a1, base);
    box_state<LOUT, S, SHIFT>(c0, c1, base);
    collide_tile<LIN, LOUT, S, R, P, SHIFT, false>(L, a0, a1, const_cast<void *>(c0), const_cast<void *>(c1),
                                                   phi + base, phio + base, xc, Ext{}, lt % gz, (lt / gz) % gy,
                                                   lt / (gz * gy));
}

// ------------------------------------------ derivatives output (per stage)
// grad phi, n and flag of the current phi over the whole box, as FP64
// (derivatives, S:498-532), for stage-level parity checks.
template <class P>
__global__ __launch_bounds__(256) void k_derivatives(Lat L, const P *__restrict__ phi, double *__restrict__ out,
                                                     int *__restrict__ flag)
{
    const int z = blockIdx.x * TZ + threadIdx.x;
    const int y = blockIdx.y * TY + threadIdx.y;
    const int x = blockIdx.z;
    if (z >= L.nz || y >= L.ny) return;
    unsigned xo[3], yo[3], zo[3];
    node_offsets(L, x, y, z, xo, yo, zo);
    using R = double;
    R gx, gy, gz;
    grad19<R>([&](int dx, int dy, int dz) { return R(phi[xo[1 + dx] + yo[1 + dy] + zo[1 + dz]]); }, gx, gy, gz);
    R fnorm = __builtin_sqrt(gx * gx + gy * gy + gz * gz);
    const size_t N = (size_t)L.N;
    const int n = xo[1] + yo[1] + zo[1];
    bool fl = fnorm > EVAP_LIM;
    out[n] = gx;
    out[N + n] = gy;
    out[2 * N + n] = gz;
    out[3 * N + n] = fl ? gx / fnorm : 0.0;
    out[4 * N + n] = fl ? gy / fnorm : 0.0;
    out[5 * N + n] = fl ? gz / fnorm : 0.0;
    flag[n] = fl ? 1 : 0;
}

// -------------------------------------------------- host-layout conversion
// Moments of a host-layout FP64 AoS array pair (post-stream populations,
// (0:18, zl, yl, xl) order): phi at each node, used as the shift reference of
// the uploaded state.
template <class P>
__global__ __launch_bounds__(256) void k_phi_aos(Lat L, const double *__restrict__ ra, const double *__restrict__ ba,
                                                 P *__restrict__ phi)
{
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= (size_t)L.N) return;
    double sr = 0, sb = 0;
    for (int i = 0; i < Q; ++i) sr += ra[n * Q + i];
    for (int i = 0; i < Q; ++i) sb += ba[n * Q + i];
    phi[n] = P((sr - sb) / (sr + sb));
}

// Pack one colour: S_i(y) = P_i(y + c_i) - w_i rho0 s(y + c_i) (inverse of
// the periodic streaming permutation, so that the next pull gather returns P;
// the shift is that of the reading node y + c_i).
// With wall faces, a slot whose reader x + c_i is a ghost node holds the
// population that bounces back into x: P_opc(i)(x).
template <class S, class P, bool SHIFT>
__global__ __launch_bounds__(256) void k_pack(Lat L, const double *__restrict__ aos, S *__restrict__ soa,
                                              const P *__restrict__ phio, int red, Ext E = Ext{})
{
    const int z = blockIdx.x * TZ + threadIdx.x;
    const int y = blockIdx.y * TY + threadIdx.y;
    const int x = blockIdx.z;
    if (z >= L.nz || y >= L.ny) return;
    unsigned xo[3], yo[3], zo[3];
    node_offsets(L, x, y, z, xo, yo, zo);
    const int n = xo[1] + yo[1] + zo[1];
    const size_t N = (size_t)L.N;
    const bool live = !is_ghost(E, L, x, y, z);
    for (int i = 0; i < Q; ++i) {
        int src = xo[1 + CX[i]] + yo[1 + CY[i]] + zo[1 + CZ[i]];
        int qi = i;
        if (live) {
            const int rx = x + CX[i], ry = y + CY[i], rz = z + CZ[i];
            if (is_wall_ghost(E, L, rx, ry, rz)) {
                src = n;
                qi = OPC[i];
            } else if (const unsigned A = mirror_axes(E, L, rx, ry, rz)) {
                // read by the reader's mirror image in the reflected direction
                src = xo[(A & 1u) ? 1 : 1 + CX[i]] + yo[(A & 2u) ? 1 : 1 + CY[i]] + zo[(A & 4u) ? 1 : 1 + CZ[i]];
                qi = REFL.d[A][i];
            }
        }
        double v = aos[(size_t)src * Q + qi];
        if constexpr (SHIFT) v = v - W[i] * (L.rho0 * shift_d(float(phio[src]), red));
        StoreOf<S>::st(soa + (size_t)i * N + n, (decltype(StoreOf<S>::ld(soa)))v);
    }
}

// Unpack one colour: P_i(x) = S_i(x - c_i) + w_i rho0 a(x), as FP64 AoS.
// With wall faces, a population that would come from a ghost node is the
// node's own opposite slot (half-way bounce-back).
template <class S, class P, bool SHIFT>
__global__ __launch_bounds__(256) void k_unpack(Lat L, const S *__restrict__ soa, double *__restrict__ aos,
                                                const P *__restrict__ phio, int red, Ext E = Ext{})
{
    const int z = blockIdx.x * TZ + threadIdx.x;
    const int y = blockIdx.y * TY + threadIdx.y;
    const int x = blockIdx.z;
    if (z >= L.nz || y >= L.ny) return;
    unsigned xo[3], yo[3], zo[3];
    node_offsets(L, x, y, z, xo, yo, zo);
    const int n = xo[1] + yo[1] + zo[1];
    const size_t N = (size_t)L.N;
    if (is_ghost(E, L, x, y, z)) {
        // ghost layers are not part of the state: no kernel writes their slots
#pragma unroll
        for (int i = 0; i < Q; ++i) aos[(size_t)n * Q + i] = 0.0;
        return;
    }
    const unsigned bb = wall_bb_mask(E, L, x, y, z);
    const unsigned madj = mirror_adjacent(E, L, x, y, z);
    static_for<Q>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        unsigned src;
        int qi;
        pull_slot<i>(xo, yo, zo, bb, madj, src, qi);
        double v = double(StoreOf<S>::ld(soa + (size_t)qi * N + src));
        if constexpr (SHIFT) v = v + W[i] * (L.rho0 * shift_d(float(phio[n]), red));
        aos[(size_t)n * Q + i] = v;
    });
}

// Unpack the factorised state: P_i(x) = r_i or b_i rebuilt at x - c_i, as
// FP64 AoS (the same reconstruction the pull gather does).
template <class S, bool SHIFT>
__global__ __launch_bounds__(256) void k_unpack_fs(Lat L, const S *__restrict__ fs, const void *__restrict__ scv,
                                                   double *__restrict__ aos, int red)
{
    const int z = blockIdx.x * TZ + threadIdx.x;
    const int y = blockIdx.y * TY + threadIdx.y;
    const int x = blockIdx.z;
    if (z >= L.nz || y >= L.ny) return;
    unsigned xo[3], yo[3], zo[3];
    node_offsets(L, x, y, z, xo, yo, zo);
    const int n = xo[1] + yo[1] + zo[1];
    const size_t N = (size_t)L.N;
    for (int i = 0; i < Q; ++i) {
        const unsigned m = xo[1 - CX[i]] + yo[1 - CY[i]] + zo[1 - CZ[i]];
        double f = double(StoreOf<S>::ld(fs + (size_t)i * N + m)), r;
        if constexpr (!SHIFT) {
            const double *sc = (const double *)scv + (size_t)m * 6;
            r = sc[0] * f;
            if (i > 0) r = r + (sc[1] * W[i]) * dot_c<double>(i, sc[2], sc[3], sc[4]);
        } else {
            const float4 g = ((const float4 *)scv)[m];
            double A, B;
            decode_frac(g.x, A, B); // exact in FP64
            f = f + W[i] * L.rho0;
            const double gd = W[i] * (CX[i] * double(g.y) + CY[i] * double(g.z) + CZ[i] * double(g.w));
            if (__builtin_signbitf(g.x)) { // red majority: the blue part first
                const double b = B * f - gd;
                r = f - b;
            } else {
                r = A * f + gd;
            }
        }
        aos[(size_t)n * Q + i] = red ? r : f - r;
    }
}

// Unpack the factorised state with walls, mirrors and particles: slot i of
// a live node from x - c_i, split with the scalars fs_src names, or with the
// side arrays (side0, side1) where those are FS_SOLID; zero on ghost layers
// (as k_unpack).
template <class S, class P, bool SHIFT>
__global__ __launch_bounds__(256) void k_unpack_fs_ext(Lat L, Ext E, const S *__restrict__ fs,
                                                       const void *__restrict__ scv, const void *__restrict__ side0,
                                                       const void *__restrict__ side1, double *__restrict__ aos,
                                                       int red)
{
    const int z = blockIdx.x * TZ + threadIdx.x;
    const int y = blockIdx.y * TY + threadIdx.y;
    const int x = blockIdx.z;
    if (z >= L.nz || y >= L.ny) return;
    unsigned xo[3], yo[3], zo[3];
    node_offsets(L, x, y, z, xo, yo, zo);
    const unsigned n = xo[1] + yo[1] + zo[1];
    const size_t N = (size_t)L.N;
    if (is_ghost(E, L, x, y, z)) {
#pragma unroll
        for (int i = 0; i < Q; ++i) aos[(size_t)n * Q + i] = 0.0;
        return;
    }
    const unsigned bb = wall_bb_mask(E, L, x, y, z);
    const unsigned madj = mirror_adjacent(E, L, x, y, z);
    static_for<Q>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        const unsigned m = xo[1 - CX[i]] + yo[1 - CY[i]] + zo[1 - CZ[i]];
        int o[3];
        unsigned flip;
        fs_src<i>(bb, madj, o, flip);
        const unsigned ms = xo[1 + o[0]] + yo[1 + o[1]] + zo[1 + o[2]];
        const size_t k = (size_t)i * N + m;
        double v;
        if constexpr (!SHIFT) {
            const FsScal64 sc = fs64_at((const double *)scv, ms);
            if (solid_scal(sc)) {
                v = double(StoreOf<S>::ld((const S *)(red ? side0 : side1) + k));
            } else {
                const double f = double(StoreOf<S>::ld(fs + k));
                double r = sc.a * f;
                if (i > 0) r = r + (sc.g * W[i]) * dot_cf<i, double>(flip, sc.nx, sc.ny, sc.nz);
                v = red ? r : f - r;
            }
        } else {
            const float4 g = ((const float4 *)scv)[ms];
            const double f = double(StoreOf<S>::ld(fs + k)) + W[i] * L.rho0;
            double A, B, r;
            if (solid_scal(g)) {
                dec_slot(((const float *)side0)[k], A, B); // exact in FP64
                r = A * f;
            } else {
                decode_frac(g.x, A, B); // exact in FP64
                const double gd = W[i] * gdot_f<i, double>(flip, g.y, g.z, g.w);
                if (__builtin_signbitf(g.x)) { // red majority: the blue part first
                    const double b = B * f - gd;
                    r = f - b;
                } else {
                    r = A * f + gd;
                }
            }
            v = red ? r : f - r;
        }
        aos[(size_t)n * Q + i] = v;
    });
}

// Wall and mirror ghost nodes of the factorised layout with walls, mirrors
// and particles: slot i of ghost node g, read by the live node x = g + c_i,
// takes the total (and the side arrays' colours) of the slot that the
// reference's boundary fills put there (S:1443-1592; pull_slot): x's own
// opposite slot off a wall (walls first), the mirror image's reflected slot
// across mirror faces.  One thread per node of face blockIdx.y (axis y / 2,
// side y % 2); edge and corner ghosts are written once per face with equal
// values.
template <class S, bool SHIFT>
__global__ __launch_bounds__(256) void k_fill_ghosts(Lat L, Ext E, S *__restrict__ fs, void *__restrict__ side0,
                                                     void *__restrict__ side1)
{
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    const int axis = blockIdx.y / 2, side = blockIdx.y % 2;
    if (!E.wall[axis][side] && !E.mirror[axis][side]) return;
    const int n3[3] = {L.nx, L.ny, L.nz};
    const int u = (axis + 1) % 3, v = (axis + 2) % 3;
    if (t >= (unsigned)(n3[u] * n3[v])) return;
    int g[3];
    g[u] = (int)(t / n3[v]);
    g[v] = (int)(t % n3[v]);
    g[axis] = side ? n3[axis] - 1 : 0;
    const size_t N = (size_t)L.N;
    const unsigned ng = ((unsigned)g[0] * L.ny + g[1]) * L.nz + g[2];
    static_for<Q>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        if (i == 0) return;
        const int x = wrapm(g[0] + CX[i], L.nx), y = wrapm(g[1] + CY[i], L.ny), z = wrapm(g[2] + CZ[i], L.nz);
        if (is_ghost(E, L, x, y, z)) return;
        unsigned xo[3], yo[3], zo[3];
        node_offsets(L, x, y, z, xo, yo, zo);
        unsigned m;
        int d;
        pull_slot<i>(xo, yo, zo, wall_bb_mask(E, L, x, y, z), mirror_adjacent(E, L, x, y, z), m, d);
        const size_t ks = (size_t)d * N + m, kd = (size_t)i * N + ng;
        fs[kd] = fs[ks];
        if constexpr (SHIFT) {
            ((float *)side0)[kd] = ((const float *)side0)[ks];
        } else {
            ((S *)side0)[kd] = ((const S *)side0)[ks];
            ((S *)side1)[kd] = ((const S *)side1)[ks];
        }
    });
}

// ------------------------------------------------- device initial condition
// Liquid cylinder along z through (xc, yc) (1-based lattice coordinates),
// radius R(z) = R0 + sum_m a_m cos(2 pi m (z - 1) / nz), tanh profile
// phi = tanh(kprof (R(z) - r)) and Laplace density jump of the reference's
// initialization (S:871-889), fluid at rest, SRT equilibria.
struct CylIC {
    double xc, yc, r0, kprof;
    int nmodes;
    int mode; // 0: liquid cylinder along z; 1: flat interface at z = z0, red below
    double amp[16];
    double z0;
};

// phi and total density of the initial state at node (x, y, z).
__device__ __forceinline__ double ic_phi(const Lat &L, const CylIC &ic, int x, int y, int z, double *rho)
{
    if (ic.mode == 1) {
        *rho = L.rho0;
        return tanh((ic.z0 - (double)(z + 1)) * ic.kprof);
    }
    const double PI = 3.14159265358979323846;
    double Rz = ic.r0;
    for (int m = 0; m < ic.nmodes; ++m) Rz += ic.amp[m] * cos(2.0 * PI * (m + 1) * (double)z / (double)L.nz);
    double dx = (double)(x + 1) - ic.xc, dy = (double)(y + 1) - ic.yc;
    const double ph = tanh((Rz - sqrt(dx * dx + dy * dy)) * ic.kprof);
    *rho = ph > 0.0 ? L.rho0 + L.sigma / Rz * 3.0 : L.rho0;
    return ph;
}

// Pass 1: phi of the initial state, the shift reference of each reader.
template <class P>
__global__ __launch_bounds__(256) void k_init_phi(Lat L, CylIC ic, P *__restrict__ phio)
{
    const int z = blockIdx.x * TZ + threadIdx.x;
    const int y = blockIdx.y * TY + threadIdx.y;
    const int x = blockIdx.z;
    if (z >= L.nz || y >= L.ny) return;
    double rho;
    phio[((size_t)x * L.ny + y) * L.nz + z] = P(ic_phi(L, ic, x, y, z, &rho));
}

// Pass 2: slot i of node s holds the equilibrium w_i rho_c of the node that
// reads it (the pull layout: s + c_i, the node itself through a wall
// bounce-back, the mirror image through a mirror face), shifted by that
// reader's fraction.
template <class S, class P, bool SHIFT>
__global__ __launch_bounds__(256) void k_init_pops(Lat L, CylIC ic, S *__restrict__ rs, S *__restrict__ bs,
                                                   const P *__restrict__ phio, Ext E = Ext{})
{
    const int z = blockIdx.x * TZ + threadIdx.x;
    const int y = blockIdx.y * TY + threadIdx.y;
    const int x = blockIdx.z;
    if (z >= L.nz || y >= L.ny) return;
    const size_t N = (size_t)L.N;
    const unsigned s = ((unsigned)x * L.ny + y) * L.nz + z;
    const bool live = !is_ghost(E, L, x, y, z);
    for (int i = 0; i < Q; ++i) {
        int r[3] = {wrap1(x + CX[i], L.nx), wrap1(y + CY[i], L.ny), wrap1(z + CZ[i], L.nz)};
        if (live) {
            if (is_wall_ghost(E, L, r[0], r[1], r[2])) {
                r[0] = x;
                r[1] = y;
                r[2] = z;
            } else if (const unsigned A = mirror_axes(E, L, r[0], r[1], r[2])) {
                if (A & 1u) r[0] = x;
                if (A & 2u) r[1] = y;
                if (A & 4u) r[2] = z;
            }
        }
        double rho;
        const double ph = ic_phi(L, ic, r[0], r[1], r[2], &rho);
        const double rr = rho * (ph + 1.0) / 2.0, rb = rho - rr;
        double vr = rr * W[i], vb = rb * W[i];
        if constexpr (SHIFT) {
            const P phs = phio[((size_t)r[0] * L.ny + r[1]) * L.nz + r[2]]; // the value pass 1 stored for the reader
            vr -= W[i] * (L.rho0 * shift_d(float(phs), 1));
            vb -= W[i] * (L.rho0 * shift_d(float(phs), 0));
        }
        StoreOf<S>::st(rs + (size_t)i * N + s, (decltype(StoreOf<S>::ld(rs)))vr);
        StoreOf<S>::st(bs + (size_t)i * N + s, (decltype(StoreOf<S>::ld(bs)))vb);
    }
}

} // namespace cgl
