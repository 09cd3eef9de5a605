// cgl_particles.hpp -- moving-particle kernels of the lattice update for
// gfx950: masks (set_walls, S:326-395), refill of uncovered nodes
// (refilling, S:1351-1418), ghost phi at live solid nodes (S:468-492),
// moving-wall bounce-back (curve, S:990-1034), momentum exchange
// (momentum, S:1178-1261), Lagrange-point capillary force (Lagrange_points
// + capillary_force2, S:1799-2028), repulsion (S:1594-1671) and the explicit
// Newton-Euler update (movement, S:1337-1346).  All particle quantities are
// FP64.
//
// The mask kernels visit the bounding box of each particle (grid.x =
// particle, grid.y = slices of the box) and k_lists then compacts, per
// particle and in box order, its live fluid neighbours (with the bit mask of
// their links into it) and its live solid nodes; the bounce-back, momentum,
// refill, wetting and ghost-phi kernels walk those lists.  Per-particle sums
// are reduced in FP64 by wavefront shuffles, then over the workgroup's waves
// and over workgroups in a fixed order, so results are reproducible from run
// to run.
//
// Positions are the reference's 1-based lattice coordinates.  On periodic
// axes a box is unwrapped around the particle, so displacements are the
// minimum-image ones.
#pragma once

#include "cgl_kernels.hpp"

namespace cgl {

constexpr int PB = 256;        // threads per particle-kernel workgroup
constexpr int PSLICES = 48;    // workgroups per particle for box kernels
constexpr int LSLICES = 16;    // slices of the link lists per particle (one workgroup each)
constexpr int LISTS_WIDE_MAX = 16; // particles up to which k_lists takes 1 024-thread workgroups
constexpr int LAG_BLOCKS = 64; // workgroups per particle for the Lagrange points (one point per thread)
constexpr int LA_M = 180, LA_N = 90, LA_NUM = LA_M * (LA_N + 1);

struct Part {
    int npar;
    int pad;
    double fg;                   // repulsion scale Fg (S:84)
    double rep_plane[6];         // repulsion planes per face (S:1607-1612)
    int rep_on[6];
    double dom_lo[3], dom_hi[3]; // Lagrange points count inside this box only (S:1832)
    int dom_lo_on[3], dom_lo_strict[3], dom_hi_on[3];
    int sym_rules;               // 1: the reference's symmetry-plane rules in movement (S:1272-1335)
    int pad2;
    double *center0;             // [npar][3] centres at set-up (the rules' plane test)
    double *center, *center_old, *center_prev; // [npar][3]: now, last step, the step before
    double *up, *rp;             // velocity, angular velocity
    double *fp, *tp, *frep, *ten, *tent; // momentum-exchange force/torque, repulsion, capillary force/torque
    double *rad, *mass, *inertia, *cita, *shb;
    double *part_lag;            // [npar][LAG_BLOCKS][7]
    double *part_mom;            // [npar][LSLICES][6]
    // link lists (k_lists) in LSLICES slices per particle, each slice the
    // nodes of one contiguous range of box_of(centre, R, 1), lcap entries per
    // slice, as box indices k: live fluid nodes next to p (pout = p+1), with
    // bit i of lmask set when x - c_i is p's solid node, and p's live solid
    // nodes (pin = p+1).  Slice s of particle p: entries from
    // ((p LSLICES + s) lcap), lcount[2 (p LSLICES + s)] and [.. + 1] of them.
    int lcap;
    int pad3;
    unsigned *lout, *lmask, *lin;
    int *lcount;
};

// Workgroup (p, s) of a list kernel walks slice s of particle p's list
// (which = 0 live fluid, 1 live solid): j over the slice's entries, at
// index base + j.
#define FOR_LIST(Pt, which, j, base)                                                                          \
    const size_t base = ((size_t)blockIdx.x * LSLICES + blockIdx.y) * (size_t)(Pt).lcap;                     \
    const int cnt_##j = (Pt).lcount[2 * (blockIdx.x * LSLICES + blockIdx.y) + (which)];                       \
    for (int j = (int)threadIdx.x; j < cnt_##j; j += PB)

// The box of particle p around centre c with margin m: unwrapped 0-based
// lower corner and extent per axis.
struct Box {
    int lo[3], w[3];
};
__device__ __forceinline__ Box box_of(const double *c, double R, double m)
{
    Box b;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        b.lo[a] = (int)floor(c[a] - 1.0 - R - m);
        int hi = (int)ceil(c[a] - 1.0 + R + m);
        b.w[a] = hi - b.lo[a] + 1;
    }
    return b;
}

// Node k of the box: unwrapped coordinates X[3] and the node index; false
// when the node falls outside a wall-bounded axis.
__device__ __forceinline__ bool box_node(const Lat &L, const Box &b, unsigned k, int X[3], unsigned &n)
{
    const int n3[3] = {L.nx, L.ny, L.nz};
    int c[3];
    c[2] = (int)(k % b.w[2]);
    k /= b.w[2];
    c[1] = (int)(k % b.w[1]);
    c[0] = (int)(k / b.w[1]);
    int g[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        X[a] = b.lo[a] + c[a];
        g[a] = wrapn(X[a], n3[a]);
    }
    n = ((unsigned)g[0] * L.ny + g[1]) * L.nz + g[2];
    return true;
}
__device__ __forceinline__ bool bounded(const Ext &E, int a)
{
    return E.wall[a][0] || E.wall[a][1] || E.mirror[a][0] || E.mirror[a][1];
}
__device__ __forceinline__ bool box_in_domain(const Ext &E, const Lat &L, const int X[3])
{
    const int n3[3] = {L.nx, L.ny, L.nz};
#pragma unroll
    for (int a = 0; a < 3; ++a)
        if (bounded(E, a) && (X[a] < 0 || X[a] >= n3[a])) return false;
    return true;
}
// not in a wall or mirror ghost layer
__device__ __forceinline__ bool interior(const Ext &E, const Lat &L, const int X[3])
{
    const int n3[3] = {L.nx, L.ny, L.nz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if ((E.wall[a][0] || E.mirror[a][0]) && X[a] < 1) return false;
        if ((E.wall[a][1] || E.mirror[a][1]) && X[a] > n3[a] - 2) return false;
    }
    return true;
}
__device__ __forceinline__ unsigned nb_index(const Lat &L, const int X[3], int dx, int dy, int dz)
{
    const int gx = wrapn(X[0] + dx, L.nx), gy = wrapn(X[1] + dy, L.ny), gz = wrapn(X[2] + dz, L.nz);
    return ((unsigned)gx * L.ny + gy) * L.nz + gz;
}
__device__ __forceinline__ unsigned box_size(const Box &b) { return (unsigned)b.w[0] * b.w[1] * b.w[2]; }

// The slot (node m, direction d) that the pull gather of live node X reads
// for direction i: the upstream node X - c_i, or, where that node lies in
// mirror ghost layers along axes A (and in no wall ghost layer), its mirror
// image (ghost coordinates replaced by X's own) in direction i with the A
// components reversed.  That is the pull form of the mirror fills followed by
// streaming (S:1443-1542, S:743-753); the moving-wall bounce-back writes, and
// the momentum exchange reads, the value there.
__device__ __forceinline__ void read_slot(const Ext &E, const Lat &L, const int X[3], int i, unsigned &m, int &d)
{
    const int n3[3] = {L.nx, L.ny, L.nz};
    const int c[3] = {CX[i], CY[i], CZ[i]};
    int src[3];
    unsigned A = 0;
    bool wall = false;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        src[a] = X[a] - c[a];
        const int g = wrapn(src[a], n3[a]);
        if ((E.wall[a][0] && g == 0) || (E.wall[a][1] && g == n3[a] - 1)) wall = true;
        if ((E.mirror[a][0] && g == 0) || (E.mirror[a][1] && g == n3[a] - 1)) A |= 1u << a;
    }
    d = i;
    if (A && !wall) {
#pragma unroll
        for (int a = 0; a < 3; ++a)
            if (A & (1u << a)) src[a] = X[a];
        d = REFL.d[A][i];
    }
    m = nb_index(L, src, 0, 0, 0);
}

// Mirror ghost nodes of face `blockIdx.y` (axis = y / 2, side = y % 2) as
// (destination, source) index pairs: the source clamps the coordinates of
// every mirror axis into the live range, which is what the reference's
// face-by-face plane copies (x, y, z; high face first) leave on edges and
// corners.  `A` gets the mirror axes the destination lies on.
__device__ __forceinline__ bool mirror_pair(const Ext &E, const Lat &L, size_t &dst, size_t &src, unsigned &A)
{
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    const int axis = blockIdx.y / 2, side = blockIdx.y % 2;
    if (!E.mirror[axis][side]) return false;
    const int n3[3] = {L.nx, L.ny, L.nz};
    const int u = (axis + 1) % 3, v = (axis + 2) % 3;
    if (t >= (unsigned)(n3[u] * n3[v])) return false;
    int cd[3], cs[3];
    cd[u] = (int)(t / n3[v]);
    cd[v] = (int)(t % n3[v]);
    cd[axis] = side ? n3[axis] - 1 : 0;
    A = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        cs[a] = cd[a];
        if (E.mirror[a][0] && cs[a] < 1) { cs[a] = 1; A |= 1u << a; }
        if (E.mirror[a][1] && cs[a] > n3[a] - 2) { cs[a] = n3[a] - 2; A |= 1u << a; }
    }
    dst = ((size_t)cd[0] * L.ny + cd[1]) * L.nz + cd[2];
    src = ((size_t)cs[0] * L.ny + cs[1]) * L.nz + cs[2];
    return true;
}

// Masks of the mirror ghost layers: copies of the adjacent layer (the
// mirror-face copies of set_walls, S:375-393).
static __global__ __launch_bounds__(256) void k_mirror_masks(Lat L, Ext E, unsigned *nt, int *walls, int *pin,
                                                             int *pout)
{
    size_t d, s;
    unsigned A;
    if (!mirror_pair(E, L, d, s, A)) return;
    nt[d] = nt[s];
    walls[d] = walls[s];
    pin[d] = pin[s];
    pout[d] = pout[s];
}

// Gradient and unit normal of the particles' band on the mirror ghost layers:
// the image's values with the components normal to the mirrors reversed
// (S:589-594), for the capillary kernel of the next step.
static __global__ __launch_bounds__(256) void k_mirror_der(Lat L, Ext E, double *der)
{
    size_t d, s;
    unsigned A;
    if (!mirror_pair(E, L, d, s, A)) return;
    const size_t N = (size_t)L.N;
#pragma unroll
    for (int k = 0; k < 6; ++k) der[k * N + d] = (A & (1u << (k % 3))) ? -der[k * N + s] : der[k * N + s];
}

// Macroscopic fields of the particles' band on the mirror ghost layers:
// densities copied, velocity with the components normal to the mirrors
// reversed (the mirror block of calmacros and refilling, S:809-828,
// S:1397-1416).
static __global__ __launch_bounds__(256) void k_mirror_mac(Lat L, Ext E, double *mac)
{
    size_t d, s;
    unsigned A;
    if (!mirror_pair(E, L, d, s, A)) return;
    const size_t N = (size_t)L.N;
#pragma unroll
    for (int k = 0; k < 3; ++k) mac[k * N + d] = mac[k * N + s];
#pragma unroll
    for (int k = 0; k < 3; ++k) mac[(3 + k) * N + d] = (A & (1u << k)) ? -mac[(3 + k) * N + s] : mac[(3 + k) * N + s];
}

// Rigid-body velocity at lattice node X (1-based position X+1) of particle
// p, from centre c (S:1009-1014).
__device__ __forceinline__ void wall_vel(const Part &P, int p, const double *c, const int X[3], double xr[3],
                                         double uw[3])
{
    xr[0] = (double)(X[0] + 1) - c[0];
    xr[1] = (double)(X[1] + 1) - c[1];
    xr[2] = (double)(X[2] + 1) - c[2];
    const double *u = P.up + 3 * p, *w = P.rp + 3 * p;
    uw[0] = u[0] + xr[2] * w[1] - xr[1] * w[2];
    uw[1] = u[1] + xr[0] * w[2] - xr[2] * w[0];
    uw[2] = u[2] + xr[1] * w[0] - xr[0] * w[1];
}

// ------------------------------------------------------------------ masks
// Zero the masks around centres c with margin m (int fields optional).
static __global__ __launch_bounds__(PB) void k_masks_clear(Lat L, Ext E, Part P, const double *ctr, double margin,
                                                    unsigned *nt, int *walls, int *pin, int *pout)
{
    const int p = blockIdx.x;
    const Box b = box_of(ctr + 3 * p, P.rad[p], margin);
    for (unsigned k = blockIdx.y * PB + threadIdx.x; k < box_size(b); k += gridDim.y * PB) {
        int X[3];
        unsigned n;
        box_node(L, b, k, X, n);
        if (!box_in_domain(E, L, X)) continue;
        nt[n] = 0;
        if (walls) {
            walls[n] = 0;
            pin[n] = 0;
            pout[n] = 0;
        }
    }
}

// Solid nodes by the FP64 distance test of S:344-345 (same operation order);
// the highest particle index wins where spheres overlap, as in the serial
// reference.  NEAR marks the band within R+3 whose macroscopic fields and
// gradients the particle kernels read.
static __global__ __launch_bounds__(PB) void k_masks_solid(Lat L, Ext E, Part P, unsigned *nt, int *walls)
{
    const int p = blockIdx.x;
    const double *c = P.center + 3 * p;
    const double R = P.rad[p];
    const Box b = box_of(c, R, 3.0);
    const double r2 = R * R, rn = (R + 3.0) * (R + 3.0);
    for (unsigned k = blockIdx.y * PB + threadIdx.x; k < box_size(b); k += gridDim.y * PB) {
        int X[3];
        unsigned n;
        box_node(L, b, k, X, n);
        if (!box_in_domain(E, L, X)) continue;
        const double dx = (double)(X[0] + 1) - c[0], dy = (double)(X[1] + 1) - c[1], dz = (double)(X[2] + 1) - c[2];
        const double dist = dx * dx + dy * dy + dz * dz;
        if (dist <= r2) {
            atomicMax(walls + n, p + 1);
            atomicOr(nt + n, NT_SOLID | NT_NEAR);
        } else if (dist <= rn) {
            atomicOr(nt + n, NT_NEAR);
        }
    }
}

// Live fluid (pout) and live solid (pin) marks (S:354-373).
static __global__ __launch_bounds__(PB) void k_masks_links(Lat L, Ext E, Part P, unsigned *nt, const int *walls, int *pin,
                                                    int *pout)
{
    const int p = blockIdx.x;
    const Box b = box_of(P.center + 3 * p, P.rad[p], 1.0);
    for (unsigned k = blockIdx.y * PB + threadIdx.x; k < box_size(b); k += gridDim.y * PB) {
        int X[3];
        unsigned n;
        box_node(L, b, k, X, n);
        if (!box_in_domain(E, L, X) || !interior(E, L, X)) continue;
        if (nt[n] & NT_SOLID) continue;
        bool link = false;
        for (int i = 1; i < Q; ++i) {
            const unsigned m = nb_index(L, X, CX[i], CY[i], CZ[i]);
            if (walls[m] == p + 1) {
                link = true;
                atomicMax(pin + m, p + 1);
                atomicOr(nt + m, NT_PIN);
            }
        }
        // the node's own marks once, however many links it has (max and or
        // give the same words in any order and number)
        if (link) {
            atomicMax(pout + n, p + 1);
            atomicOr(nt + n, NT_POUT);
        }
    }
}

// ------------------------------------------------------------------ refill
// Nodes uncovered by the move (live fluid now, solid before): rigid-body
// velocity about the old centre, unweighted neighbour means of the colour
// densities over previously fluid neighbours, equilibria in K2 (S:1364-1391).
template <class P>
__global__ __launch_bounds__(PB) void k_refill(Lat L, Ext E, Part Pt, const unsigned *nt_old, unsigned *nt_new,
                                               const int *pout, P *phi)
{
    const int p = blockIdx.x;
    const Box b = box_of(Pt.center + 3 * p, Pt.rad[p], 1.0);
    const size_t N = (size_t)L.N;
    FOR_LIST(Pt, 0, j, lb)
    {
        int X[3];
        unsigned n;
        box_node(L, b, Pt.lout[lb + j], X, n);
        if (!(nt_old[n] & NT_SOLID)) continue;
        double xr[3], uw[3];
        wall_vel(Pt, p, Pt.center_old + 3 * p, X, xr, uw);
        double sr = 0.0, sb = 0.0, num = 0.0;
        for (int i = 1; i < Q; ++i) {
            const unsigned m = nb_index(L, X, CX[i], CY[i], CZ[i]);
            if (!(nt_old[m] & NT_SOLID)) {
                sr = sr + E.mac[N + m];
                sb = sb + E.mac[2 * N + m];
                num = num + 1.0;
            }
        }
        const double rr = sr / num, rb = sb / num, rho = rr + rb;
        E.mac[n] = rho;
        E.mac[N + n] = rr;
        E.mac[2 * N + n] = rb;
        E.mac[3 * N + n] = uw[0];
        E.mac[4 * N + n] = uw[1];
        E.mac[5 * N + n] = uw[2];
        phi[n] = P((rr - rb) / rho);
        atomicOr(nt_new + n, NT_REFILL);
    }
}

// Ghost phi at live solid nodes: w-weighted mean over fluid neighbours
// (S:468-491).
template <class P>
__global__ __launch_bounds__(PB) void k_ghost_phi(Lat L, Ext E, Part Pt, const unsigned *nt, const int *pin, P *phi)
{
    const int p = blockIdx.x;
    const Box b = box_of(Pt.center + 3 * p, Pt.rad[p], 1.0);
    FOR_LIST(Pt, 1, j, lb)
    {
        int X[3];
        unsigned n;
        box_node(L, b, Pt.lin[lb + j], X, n);
        double wsum = 0.0, acc = 0.0;
        for (int i = 1; i < Q; ++i) {
            const unsigned m = nb_index(L, X, CX[i], CY[i], CZ[i]);
            if (!(nt[m] & NT_SOLID)) {
                wsum = wsum + W[i];
                acc = acc + W[i] * double(phi[m]);
            }
        }
        phi[n] = P(acc / wsum);
    }
}

// Link lists of particle p, slice s = blockIdx.y: the box nodes from
// s * per to (s + 1) * per (per = ceil(box / LSLICES)) in box order, each
// chunk of PB of them taking its place by a block-wide prefix (wave ballots,
// then the waves in order), so the lists, and every sum over them, are the
// same from run to run.  LB threads per workgroup: a slice of a box is a few
// hundred to a thousand nodes, taken in chunks of LB one after another, each
// a chain of dependent loads; with few particles (LSLICES workgroups each)
// nothing else overlaps them, so such a handle takes 1 024-thread chunks
// (33 -> 11 us per step for one sphere of radius 10), while 200 particles
// keep 256 (1 024 idles most lanes of a slice there: 47 -> 110 us).
template <int LB>
__global__ __launch_bounds__(LB) void k_lists(Lat L, Ext E, Part P, const unsigned *nt, const int *walls,
                                                     const int *pin, const int *pout)
{
    const int p = blockIdx.x;
    const Box b = box_of(P.center + 3 * p, P.rad[p], 1.0);
    const unsigned nbox = box_size(b), per = (nbox + LSLICES - 1) / LSLICES;
    const unsigned k1 = min(nbox, (blockIdx.y + 1) * per);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __shared__ unsigned wsum[2][LB / 64];
    __shared__ unsigned base[2];
    if (threadIdx.x == 0) base[0] = base[1] = 0;
    __syncthreads();
    const size_t off = ((size_t)p * LSLICES + blockIdx.y) * (size_t)P.lcap;
    for (unsigned k0 = blockIdx.y * per; k0 < k1; k0 += LB) {
        const unsigned k = k0 + threadIdx.x;
        bool isout = false, isin = false;
        unsigned mask = 0;
        if (k < k1) {
            int X[3];
            unsigned n;
            box_node(L, b, k, X, n);
            if (box_in_domain(E, L, X) && interior(E, L, X)) {
                if (pout[n] == p + 1 && !(nt[n] & NT_SOLID)) {
                    isout = true;
                    for (int i = 1; i < Q; ++i)
                        if (walls[nb_index(L, X, -CX[i], -CY[i], -CZ[i])] == p + 1) mask |= 1u << i;
                }
                isin = pin[n] == p + 1;
            }
        }
        const unsigned long long bo = __ballot(isout), bi = __ballot(isin);
        const unsigned long long below = (1ull << lane) - 1ull;
        if (lane == 0) {
            wsum[0][wave] = (unsigned)__popcll(bo);
            wsum[1][wave] = (unsigned)__popcll(bi);
        }
        __syncthreads();
        unsigned jo = base[0], ji = base[1];
        for (int w = 0; w < wave; ++w) {
            jo += wsum[0][w];
            ji += wsum[1][w];
        }
        jo += (unsigned)__popcll(bo & below);
        ji += (unsigned)__popcll(bi & below);
        if (isout && jo < (unsigned)P.lcap) {
            P.lout[off + jo] = k;
            P.lmask[off + jo] = mask;
        }
        if (isin && ji < (unsigned)P.lcap) P.lin[off + ji] = k;
        __syncthreads();
        if (threadIdx.x == 0)
            for (int w = 0; w < LB / 64; ++w) {
                base[0] += wsum[0][w];
                base[1] += wsum[1][w];
            }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const size_t c = 2 * ((size_t)p * LSLICES + blockIdx.y);
        P.lcount[c] = (int)min(base[0], (unsigned)P.lcap);
        P.lcount[c + 1] = (int)min(base[1], (unsigned)P.lcap);
    }
}

// --------------------------------------------------------------- bounce-back
// Wetted unit normals at the live fluid nodes next to particle p (pout = p+1,
// S:536-587), from the phase field the collide kernel reads this step, into
// the normal slots of the band fields (E.der rows 3-5); the collide kernel
// takes them from there (a mirror ghost from its image) instead of applying
// the correction itself.  Same gradient and normal arithmetic as the collide
// kernel's normals pass; zero without the interface flag.  The collide
// kernel writes the same values into these slots at band nodes.
template <class R, class P>
__global__ __launch_bounds__(PB) void k_wet_normals(Lat L, Ext E, Part Pt, const P *__restrict__ phi)
{
    const int p = blockIdx.x;
    const Box b = box_of(Pt.center + 3 * p, Pt.rad[p], 1.0);
    const size_t N = (size_t)L.N;
    FOR_LIST(Pt, 0, j, lb)
    {
        int X[3];
        unsigned n;
        box_node(L, b, Pt.lout[lb + j], X, n);
        R gx, gy, gz, nx, ny, nz;
        grad19<R>([&](int dx, int dy, int dz) { return phi[nb_index(L, X, dx, dy, dz)]; }, gx, gy, gz);
        if (unit_normal<R>(gx, gy, gz, nx, ny, nz)) {
            // the node's position in the box (wetting takes the minimum image)
            double nv[3] = {double(nx), double(ny), double(nz)};
            wetting(L, E, wrapn(X[0], L.nx) + 1.0, wrapn(X[1], L.ny) + 1.0, wrapn(X[2], L.nz) + 1.0,
                    Pt.center + 3 * p, Pt.cita[p], nv);
            nx = R(nv[0]);
            ny = R(nv[1]);
            nz = R(nv[2]);
        }
        E.der[3 * N + n] = double(nx);
        E.der[4 * N + n] = double(ny);
        E.der[5 * N + n] = double(nz);
    }
}

// Moving-wall half-way bounce-back on links from live fluid x into its
// particle (S:1004-1018):  r_i(x) = r*_opp(x) + 6 w_i rho_R(x) (u_w . c_i).
// The pull gather of the next step reads direction i of x from the solid
// node x - c_i (or, across a mirror face, from that node's image), so the
// value is written into that slot.
template <class S, bool SHIFT>
__global__ __launch_bounds__(PB) void k_curve(Lat L, Ext E, Part Pt, const unsigned *nt, const int *walls,
                                              const int *pout, S *ro, S *bo, const float *phis)
{
    const int p = blockIdx.x;
    const Box b = box_of(Pt.center + 3 * p, Pt.rad[p], 1.0);
    const size_t N = (size_t)L.N;
    FOR_LIST(Pt, 0, j, lb)
    {
        int X[3];
        unsigned n;
        box_node(L, b, Pt.lout[lb + j], X, n);
        const unsigned links = Pt.lmask[lb + j];
        double xr[3], uw[3];
        wall_vel(Pt, p, Pt.center + 3 * p, X, xr, uw);
        const double rr = E.mac[N + n], rbv = E.mac[2 * N + n];
        for (int i = 1; i < Q; ++i) {
            if (!((links >> i) & 1u)) continue;
            const unsigned m = nb_index(L, X, -CX[i], -CY[i], -CZ[i]);
            unsigned ms;
            int ds;
            read_slot(E, L, X, i, ms, ds);
            const double star = uw[0] * CX[i] + uw[1] * CY[i] + uw[2] * CZ[i];
            const int o = OPC[i];
            double vr = double(StoreOf<S>::ld(ro + (size_t)o * N, n));
            double vb = double(StoreOf<S>::ld(bo + (size_t)o * N, n));
            if constexpr (SHIFT) {
                // slot o of x is read by x + c_o = m; slot i of m is read by x
                vr += W[o] * L.rho0 * shift_d(phis[m], 1);
                vb += W[o] * L.rho0 * shift_d(phis[m], 0);
                vr = vr + 6.0 * W[i] * rr * star - W[i] * L.rho0 * shift_d(phis[n], 1);
                vb = vb + 6.0 * W[i] * rbv * star - W[i] * L.rho0 * shift_d(phis[n], 0);
            } else {
                vr = vr + 6.0 * W[i] * rr * star;
                vb = vb + 6.0 * W[i] * rbv * star;
            }
            StoreOf<S>::st(ro + (size_t)ds * N, ms, (decltype(StoreOf<S>::ld(ro)))vr);
            StoreOf<S>::st(bo + (size_t)ds * N, ms, (decltype(StoreOf<S>::ld(bo)))vb);
        }
    }
}

// k_curve for the factorised layout: the node's own post-collision slot opp
// is rebuilt from its recolour scalars (fo, so: the step's output, FP64
// arithmetic), the bounce-back terms are added per colour, and the slot that
// x gathers for direction i -- slot i of node x - c_i, the solid node or, for
// a particle across a mirror face, the ghost node that the gather reads --
// takes the total and the colours (side arrays: FP64 red and blue, shifted
// storage the slot's minority fraction).  Runs after k_fill_ghosts, whose
// value of that ghost slot it replaces, as curve follows the boundary fills
// (S:247-251).
// own (the scatter form, shifted storage): the slot's colours are also added
// to x's fixed-point phase-field sums (total from the stored value, minority
// colour formed first); a slot whose upstream node is a wall ghost is left
// to the wall's bounce-back, whose terms x added itself.
template <class S, bool SHIFT>
__global__ __launch_bounds__(PB) void k_curve_fs(Lat L, Ext E, Part Pt, S *fo, const void *so, void *side0,
                                                 void *side1, int2 *own = nullptr)
{
    const int p = blockIdx.x;
    const Box b = box_of(Pt.center + 3 * p, Pt.rad[p], 1.0);
    const size_t N = (size_t)L.N;
    FOR_LIST(Pt, 0, j, lb)
    {
        int X[3];
        unsigned n;
        box_node(L, b, Pt.lout[lb + j], X, n);
        const unsigned links = Pt.lmask[lb + j];
        double xr[3], uw[3];
        wall_vel(Pt, p, Pt.center + 3 * p, X, xr, uw);
        const double rr = E.mac[N + n], rbv = E.mac[2 * N + n];
        for (int i = 1; i < Q; ++i) {
            if (!((links >> i) & 1u)) continue;
            const unsigned m = nb_index(L, X, -CX[i], -CY[i], -CZ[i]);
            const double star = uw[0] * CX[i] + uw[1] * CY[i] + uw[2] * CZ[i];
            const int o = OPC[i];
            double f = double(StoreOf<S>::ld(fo + (size_t)o * N, n));
            double vr, vb;
            if constexpr (SHIFT) {
                const float4 g = ((const float4 *)so)[n];
                double A, B;
                decode_frac(g.x, A, B);
                const double gc = W[o] * (CX[o] * double(g.y) + CY[o] * double(g.z) + CZ[o] * double(g.w));
                f = f + W[o] * L.rho0;
                if (__builtin_signbitf(g.x)) {
                    vb = B * f - gc;
                    vr = f - vb;
                } else {
                    vr = A * f + gc;
                    vb = f - vr;
                }
            } else {
                const FsScal64 sc = fs64_at((const double *)so, n);
                vr = sc.a * f;
                vr = vr + (sc.g * W[o]) * dot_c<double>(o, sc.nx, sc.ny, sc.nz);
                vb = f - vr;
            }
            vr = vr + 6.0 * W[i] * rr * star;
            vb = vb + 6.0 * W[i] * rbv * star;
            const size_t k = (size_t)i * N + m;
            if constexpr (SHIFT) {
                fo[k] = (S)(float)(vr + vb - W[i] * L.rho0);
                ((float *)side0)[k] = enc_slot(vr, vb);
                if (own && !is_wall_ghost(E, L, wrapn(X[0] - CX[i], L.nx), wrapn(X[1] - CY[i], L.ny),
                                          wrapn(X[2] - CZ[i], L.nz))) {
                    const float qs = SC_ONE / float(L.rho0);
                    const bool redmaj = vb <= vr;
                    const int qf = round_half_up(__builtin_fmaf(float(fo[k]), qs, float(W[i] * double(SC_ONE))));
                    const int qm = round_half_up(float(redmaj ? vb : vr) * qs);
                    // one 64-bit addition, exact while the red half stays in
                    // [0, 2^32) (the own red sums carry a bias of 2^30)
                    const int cr = redmaj ? qf - qm : qm, cb = redmaj ? qm : qf - qm;
                    atomicAdd((unsigned long long *)&own[n],
                              ((unsigned long long)(unsigned)cb << 32) + (unsigned long long)(long long)cr);
                }
            } else {
                fo[k] = vr + vb;
                ((S *)side0)[k] = vr;
                ((S *)side1)[k] = vb;
            }
        }
    }
}

// ------------------------------------- special nodes of the factorised layout
// With walls, mirrors or particles in the factorised layout the tile kernels
// leave the special lanes (special_lane: ghost layers, nodes next to a ghost
// layer, the particles' band) to these per-node kernels, which read the phase
// field and the recolour scalars from memory instead of the LDS rings and
// take the general gather (bounce-back and mirror slots, side arrays of
// solid nodes).  BAND = false: one thread per node of the live layer next to
// face blockIdx.y (axis y / 2, side y % 2) where that face is a mirror;
// BAND = true: one thread per node of particle blockIdx.y's box around its
// band (R + 3), node blockIdx.x PB + threadIdx.x in box order (band_blocks
// workgroups per particle).  A node reached twice (an edge, a band next to a
// face, overlapping bands) is computed twice with equal results.
template <bool BAND, class F>
__device__ __forceinline__ void for_special_nodes(const Lat &L, const Ext &E, const Part &Pt, F f)
{
    const int n3[3] = {L.nx, L.ny, L.nz};
    if constexpr (!BAND) {
        const int axis = blockIdx.y / 2, side = blockIdx.y % 2;
        if (!E.mirror[axis][side]) return;
        const int u = (axis + 1) % 3, v = (axis + 2) % 3;
        const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
        if (t >= (unsigned)(n3[u] * n3[v])) return;
        int c[3];
        c[u] = (int)(t / n3[v]);
        c[v] = (int)(t % n3[v]);
        c[axis] = side ? n3[axis] - 2 : 1;
        f(c[0], c[1], c[2]);
    } else {
        const int p = blockIdx.y;
        const Box b = box_of(Pt.center + 3 * p, Pt.rad[p], 3.0);
        const unsigned k = blockIdx.x * PB + threadIdx.x;
        if (k >= box_size(b)) return;
        int X[3];
        unsigned n;
        box_node(L, b, k, X, n);
        if (!box_in_domain(E, L, X)) return;
        f(wrapn(X[0], L.nx), wrapn(X[1], L.ny), wrapn(X[2], L.nz));
    }
}

// The unit normal at node (px, py, pz) as the tile kernel's normals pass
// forms it with walls and particles (S:498-532, S:536-594): zero at wall
// ghosts and solid nodes, the image's reflected at mirror ghosts, the wetted
// normal stored by k_wet_normals at live fluid nodes next to a particle.
template <class R, class P>
__device__ __forceinline__ void ext_normal(const Lat &L, const Ext &E, const P *__restrict__ phi, int px, int py,
                                           int pz, R nv[3])
{
    const int gxw = wrapn(px, L.nx), gyw = wrapn(py, L.ny), gzw = wrapn(pz, L.nz);
    R nx = 0, ny = 0, nz = 0;
    bool skip = is_wall_ghost(E, L, gxw, gyw, gzw);
    unsigned nt = 0, A = 0;
    int ox = 0, oy = 0, oz = 0;
    if (!skip) {
        A = mirror_axes(E, L, gxw, gyw, gzw);
        ox = (A & 1u) ? (gxw == 0 ? 1 : -1) : 0;
        oy = (A & 2u) ? (gyw == 0 ? 1 : -1) : 0;
        oz = (A & 4u) ? (gzw == 0 ? 1 : -1) : 0;
    }
    const unsigned sx = (unsigned)L.ny * (unsigned)L.nz;
    if (!skip && E.ntype) {
        nt = at(E.ntype, (unsigned)gxw * sx + (unsigned)gyw * L.nz + gzw);
        skip = (nt & NT_SOLID) != 0;
    }
    if (!skip) {
        R gx, gy, gz;
        grad19<R>(
            [&](int dx, int dy, int dz) {
                return phi[(unsigned)wrapn(gxw + ox + dx, L.nx) * sx + (unsigned)wrapn(gyw + oy + dy, L.ny) * L.nz +
                           wrapn(gzw + oz + dz, L.nz)];
            },
            gx, gy, gz);
        const bool fl = unit_normal<R>(gx, gy, gz, nx, ny, nz);
        if (fl && (nt & NT_POUT)) {
            const size_t N = (size_t)L.N;
            const unsigned im = (unsigned)(gxw + ox) * sx + (unsigned)(gyw + oy) * L.nz + (gzw + oz);
            nx = R(E.der[3 * N + im]);
            ny = R(E.der[4 * N + im]);
            nz = R(E.der[5 * N + im]);
        }
        if (A & 1u) nx = -nx;
        if (A & 2u) ny = -ny;
        if (A & 4u) nz = -nz;
    }
    nv[0] = nx;
    nv[1] = ny;
    nv[2] = nz;
}

// K1 of the special nodes: calmacros as node_moments forms it (general
// gather), the band's FP64 fields and phi.
template <class S, class R, class P, bool SHIFT, bool BAND>
__global__ __launch_bounds__(256) void k_moments_special(Lat L, const void *__restrict__ p0,
                                                         const void *__restrict__ p1, const P *__restrict__ phio,
                                                         P *__restrict__ phi, Ext E, Part Pt)
{
    for_special_nodes<BAND>(L, E, Pt, [&](int x, int y, int z) {
        const unsigned n0 = ((unsigned)x * L.ny + y) * L.nz + z;
        if (!special_lane(E, L, x, y, z, n0)) return;
        unsigned n, nt;
        Moments<R, SHIFT> mo;
        if (!node_moments<FS, S, R, P, SHIFT, true>(L, p0, p1, phio, E, x, y, z, n, mo, nt)) return;
        band_macros(L, E, n, nt, mo);
        at(phi, n) = (nt & NT_SOLID) ? P(0.0f) : phi_entry<P>(mo);
    });
}

// K2 of the special nodes: lane_step with the phase field, the normals of
// the node and its 18 neighbours (ext_normal) and the recolour scalars read
// from memory.
template <class S, class R, class P, bool SHIFT, bool BAND>
__global__ __launch_bounds__(256) void k_collide_special(Lat L, const void *__restrict__ i0,
                                                         const void *__restrict__ i1, void *__restrict__ o0,
                                                         void *__restrict__ o1, const P *__restrict__ phi,
                                                         const P *__restrict__ phio, Ext E, Part Pt)
{
    for_special_nodes<BAND>(L, E, Pt, [&](int x, int y, int z) {
        const unsigned sx = (unsigned)L.ny * (unsigned)L.nz;
        if (!special_lane(E, L, x, y, z, (unsigned)x * sx + (unsigned)y * L.nz + z)) return;
        // normals by offset key (dx + 1) 9 + (dy + 1) 3 + dz + 1, each formed
        // at its first use: the node's own always, the neighbours' only for
        // the curvature of an interface node (the keys are compile-time
        // constants in lane_step, so the flags fold away)
        R nn[27][3];
        bool have[27] = {};
        lane_step<FS, FS, S, R, P, SHIFT, true>(
            L, i0, i1, o0, o1, phio, E, x, y, z,
            [&](int dx, int dy, int dz) {
                return phi[(unsigned)wrapn(x + dx, L.nx) * sx + (unsigned)wrapn(y + dy, L.ny) * L.nz +
                           wrapn(z + dz, L.nz)];
            },
            [&](int k, int dx, int dy, int dz) {
                const int key = (dx + 1) * 9 + (dy + 1) * 3 + dz + 1;
                if (!have[key]) {
                    ext_normal<R, P>(L, E, phi, x + dx, y + dy, z + dz, nn[key]);
                    have[key] = true;
                }
                return nn[key][k];
            },
            [&](const int o[3]) {
                const unsigned m = (unsigned)wrapn(x + o[0], L.nx) * sx + (unsigned)wrapn(y + o[1], L.ny) * L.nz +
                                   wrapn(z + o[2], L.nz);
                if constexpr (SHIFT)
                    return at((const float4 *)i1, m);
                else
                    return fs64_at((const double *)i1, m);
            });
    });
}

// Fixed-order FP64 sum of K values per thread over the workgroup: an xor
// butterfly of wavefront shuffles (every lane ends with its wave's sum; the
// pairing is the same every run, and addition commutes, so all lanes agree),
// then the PB / 64 wave sums in wave order.
template <int K> __device__ __forceinline__ void block_sum(double v[K], double *out)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int k = 0; k < K; ++k) v[k] += __shfl_xor(v[k], off);
    __shared__ double ws[PB / 64][K];
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < K; ++k) ws[threadIdx.x >> 6][k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int k = 0; k < K; ++k) {
            double t = ws[0][k];
            for (int w = 1; w < PB / 64; ++w) t += ws[w][k];
            out[k] = t;
        }
}

// ---------------------------------------------------------- momentum exchange
// Galilean-invariant momentum exchange on the links of particle p
// (S:1198-1252): f_i(x) is the bounced-back population (held in the slot the
// gather of x reads), ffi_opp(x) the node's post-collision one.  On the first
// step after the start-up (cold != 0) there is no previous collision: the
// reference's initialization sets ffi = fi (S:894), the node's own
// populations, which the pull layout holds where x gathers them.
template <class S, bool SHIFT>
__global__ __launch_bounds__(PB) void k_momentum(Lat L, Ext E, Part Pt, const unsigned *nt, const int *walls,
                                                 const int *pout, const S *rs, const S *bs, int cold)
{
    const int p = blockIdx.x;
    const Box b = box_of(Pt.center + 3 * p, Pt.rad[p], 1.0);
    const size_t N = (size_t)L.N;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    FOR_LIST(Pt, 0, j, lb)
    {
        int X[3];
        unsigned n;
        box_node(L, b, Pt.lout[lb + j], X, n);
        const unsigned links = Pt.lmask[lb + j];
        double xr[3], uw[3];
        wall_vel(Pt, p, Pt.center + 3 * p, X, xr, uw);
        for (int i = 1; i < Q; ++i) {
            if (!((links >> i) & 1u)) continue;
            const int o = OPC[i];
            unsigned mi, no = n;
            int di, dof = o;
            read_slot(E, L, X, i, mi, di);
            if (cold) read_slot(E, L, X, o, no, dof);
            double fi = double(StoreOf<S>::ld(rs + (size_t)di * N, mi)) + double(StoreOf<S>::ld(bs + (size_t)di * N, mi));
            double fo = double(StoreOf<S>::ld(rs + (size_t)dof * N, no)) + double(StoreOf<S>::ld(bs + (size_t)dof * N, no));
            if constexpr (SHIFT) {
                fi += W[i] * L.rho0;
                fo += W[o] * L.rho0;
            }
            const double fmx = ((double)CX[i] - uw[0]) * fi - ((double)CX[o] - uw[0]) * fo;
            const double fmy = ((double)CY[i] - uw[1]) * fi - ((double)CY[o] - uw[1]) * fo;
            const double fmz = ((double)CZ[i] - uw[2]) * fi - ((double)CZ[o] - uw[2]) * fo;
            acc[0] -= fmx;
            acc[1] -= fmy;
            acc[2] -= fmz;
            acc[3] -= xr[1] * fmz - xr[2] * fmy;
            acc[4] -= xr[2] * fmx - xr[0] * fmz;
            acc[5] -= xr[0] * fmy - xr[1] * fmx;
        }
    }
    block_sum<6>(acc, Pt.part_mom + ((size_t)p * LSLICES + blockIdx.y) * 6);
}

// k_momentum for the factorised layout: every node's slots hold their totals
// (solid nodes' and ghost nodes' too), so f_i(x) is slot i of x - c_i, the
// bounced-back value, and ffi_opp(x) slot opp of x, the node's own
// post-collision population; at the cold start ffi is the slot x gathers
// for opp, slot opp of x - c_opp.
template <class S, bool SHIFT>
__global__ __launch_bounds__(PB) void k_momentum_fs(Lat L, Ext E, Part Pt, const S *fs, int cold)
{
    const int p = blockIdx.x;
    const Box b = box_of(Pt.center + 3 * p, Pt.rad[p], 1.0);
    const size_t N = (size_t)L.N;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    FOR_LIST(Pt, 0, j, lb)
    {
        int X[3];
        unsigned n;
        box_node(L, b, Pt.lout[lb + j], X, n);
        const unsigned links = Pt.lmask[lb + j];
        double xr[3], uw[3];
        wall_vel(Pt, p, Pt.center + 3 * p, X, xr, uw);
        for (int i = 1; i < Q; ++i) {
            if (!((links >> i) & 1u)) continue;
            const int o = OPC[i];
            const unsigned mi = nb_index(L, X, -CX[i], -CY[i], -CZ[i]);
            const unsigned no = cold ? nb_index(L, X, -CX[o], -CY[o], -CZ[o]) : n;
            double fi = double(StoreOf<S>::ld(fs + (size_t)i * N + mi));
            double fo = double(StoreOf<S>::ld(fs + (size_t)o * N + no));
            if constexpr (SHIFT) {
                fi += W[i] * L.rho0;
                fo += W[o] * L.rho0;
            }
            const double fmx = ((double)CX[i] - uw[0]) * fi - ((double)CX[o] - uw[0]) * fo;
            const double fmy = ((double)CY[i] - uw[1]) * fi - ((double)CY[o] - uw[1]) * fo;
            const double fmz = ((double)CZ[i] - uw[2]) * fi - ((double)CZ[o] - uw[2]) * fo;
            acc[0] -= fmx;
            acc[1] -= fmy;
            acc[2] -= fmz;
            acc[3] -= xr[1] * fmz - xr[2] * fmy;
            acc[4] -= xr[2] * fmx - xr[0] * fmz;
            acc[5] -= xr[0] * fmy - xr[1] * fmx;
        }
    }
    block_sum<6>(acc, Pt.part_mom + ((size_t)p * LSLICES + blockIdx.y) * 6);
}

// --------------------------------------------------- capillary force (FP64)
__device__ __forceinline__ void unitize3(double a[3])
{
    double st = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    a[0] /= st;
    a[1] /= st;
    a[2] /= st;
}
__device__ __forceinline__ void cross3(const double a[3], const double b[3], double c[3])
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = -(a[0] * b[2] - a[2] * b[0]);
    c[2] = a[0] * b[1] - a[1] * b[0];
}

// 180 x 91 surface points of sphere p (S:1821-1962) and their capillary
// force and torque (S:1993-2024): IDW-interpolated phase field over the
// fluid corners of the enclosing cell, the nearest interface corner's normal
// with the wetting correction, force 4.5 k beta (1-phi^2)^2 dS |n x n_w| sigma.
template <class P>
__global__ __launch_bounds__(PB) void k_lagrange(Lat L, Ext E, Part Pt, const P *phi, const int *walls)
{
    const int p = blockIdx.x;
    const double PI = 3.14159265358979323846;
    const double alpha = 2.0 / 180.0 * PI;
    const double kcap = 0.134 * 1.43;
    const double *c = Pt.center + 3 * p;
    const double R = Pt.rad[p], cita = Pt.cita[p];
    const size_t N = (size_t)L.N;
    static const int ex[8] = {0, 1, 1, 0, 0, 1, 1, 0};
    static const int ey[8] = {0, 0, 1, 1, 0, 0, 1, 1};
    static const int ez[8] = {0, 0, 0, 0, 1, 1, 1, 1};
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int num = blockIdx.y * PB + threadIdx.x; num < LA_NUM; num += gridDim.y * PB) {
        const int na = num / (LA_N + 1), nb = num - na * (LA_N + 1);
        const double aa = (double)na * alpha, bb = (double)nb * alpha;
        double Lp[3];
        Lp[0] = c[0] + R * sin(bb) * cos(aa);
        Lp[1] = c[1] + R * sin(bb) * sin(aa);
        Lp[2] = c[2] + R * cos(bb);
        const double ds = R * R * sin(bb) * (alpha * alpha);
        // points outside the in-domain box neither interpolate nor count
        // (S:1832): lrho = 0, no interface flag
        bool in = true;
        for (int a = 0; a < 3; ++a) {
            if (Pt.dom_lo_on[a] && (Pt.dom_lo_strict[a] ? !(Lp[a] > Pt.dom_lo[a]) : !(Lp[a] >= Pt.dom_lo[a]))) in = false;
            if (Pt.dom_hi_on[a] && !(Lp[a] < Pt.dom_hi[a])) in = false;
        }
        if (!in) continue;
        int c0[3];
        for (int a = 0; a < 3; ++a) c0[a] = (int)Lp[a];
        double wsum = 0.0, lrho = 0.0;
        for (int k = 0; k < 8; ++k) {
            const int f1[3] = {c0[0] + ex[k], c0[1] + ey[k], c0[2] + ez[k]};
            const int Xc[3] = {f1[0] - 1, f1[1] - 1, f1[2] - 1};
            const unsigned m = nb_index(L, Xc, 0, 0, 0);
            if (walls[m] <= 0) {
                const double star = double(phi[m]) + E.der[m] * (Lp[0] - (double)f1[0]) +
                                    E.der[N + m] * (Lp[1] - (double)f1[1]) + E.der[2 * N + m] * (Lp[2] - (double)f1[2]);
                const double dd = ((double)f1[0] - Lp[0]) * ((double)f1[0] - Lp[0]) +
                                  ((double)f1[1] - Lp[1]) * ((double)f1[1] - Lp[1]) +
                                  ((double)f1[2] - Lp[2]) * ((double)f1[2] - Lp[2]);
                const double wgt = 1.0 / dd;
                lrho = lrho + wgt * star;
                wsum = wsum + wgt;
            }
        }
        lrho = lrho / wsum;
        if (lrho < -1.0) lrho = -1.0;
        if (lrho > 1.0) lrho = 1.0;
        if (lrho > 0.0) acc[6] += ds;
        if (!(fabs(lrho) < 0.99)) continue;
        double nw[3] = {Lp[0] - c[0], Lp[1] - c[1], Lp[2] - c[2]};
        unitize3(nw);
        double dmin = 100.0;
        int best[3] = {c0[0], c0[1], c0[2]};
        for (int k = 0; k < 8; ++k) {
            const int f1[3] = {c0[0] + ex[k], c0[1] + ey[k], c0[2] + ez[k]};
            const int Xc[3] = {f1[0] - 1, f1[1] - 1, f1[2] - 1};
            const unsigned m = nb_index(L, Xc, 0, 0, 0);
            if (walls[m] > 0) continue;
            const double nx = E.der[3 * N + m], ny = E.der[4 * N + m], nz = E.der[5 * N + m];
            if (!(fabs(nx * nx + ny * ny + nz * nz) > 0.0)) continue;
            const double dd = ((double)f1[0] - Lp[0]) * ((double)f1[0] - Lp[0]) +
                              ((double)f1[1] - Lp[1]) * ((double)f1[1] - Lp[1]) +
                              ((double)f1[2] - Lp[2]) * ((double)f1[2] - Lp[2]);
            if (sqrt(dd) < dmin) {
                dmin = sqrt(dd);
                best[0] = f1[0];
                best[1] = f1[1];
                best[2] = f1[2];
            }
        }
        const int Xb[3] = {best[0] - 1, best[1] - 1, best[2] - 1};
        const unsigned mb = nb_index(L, Xb, 0, 0, 0);
        double nt[3] = {E.der[3 * N + mb], E.der[4 * N + mb], E.der[5 * N + mb]};
        unitize3(nt);
        const double theta = acos(nw[0] * nt[0] + nw[1] * nt[1] + nw[2] * nt[2]);
        const double ca = cos(cita), sa = sin(cita);
        const double a1 = ca - sa * cos(theta) / sin(theta), a2 = ca + sa * cos(theta) / sin(theta);
        const double b1 = sa / sin(theta), b2 = -sa / sin(theta);
        double v1[3], v2[3];
        for (int k = 0; k < 3; ++k) {
            v1[k] = a1 * nw[k] + b1 * nt[k];
            v2[k] = a2 * nw[k] + b2 * nt[k];
        }
        const double db1 = (v1[0] - nt[0]) * (v1[0] - nt[0]) + (v1[1] - nt[1]) * (v1[1] - nt[1]) +
                           (v1[2] - nt[2]) * (v1[2] - nt[2]);
        const double db2 = (v2[0] - nt[0]) * (v2[0] - nt[0]) + (v2[1] - nt[1]) * (v2[1] - nt[1]) +
                           (v2[2] - nt[2]) * (v2[2] - nt[2]);
        double ln[3];
        for (int k = 0; k < 3; ++k) ln[k] = db1 < db2 ? v1[k] : v2[k];
        // capillary_force2 (S:1996-2022)
        const double xr[3] = {Lp[0] - c[0], Lp[1] - c[1], Lp[2] - c[2]};
        double nw2[3] = {xr[0], xr[1], xr[2]};
        unitize3(nw2);
        const double ie[3] = {-ln[0], -ln[1], -ln[2]};
        double tmp[3], ntan[3];
        cross3(ie, nw2, tmp);
        const double proj = sqrt(tmp[0] * tmp[0] + tmp[1] * tmp[1] + tmp[2] * tmp[2]);
        cross3(tmp, ie, ntan);
        unitize3(ntan);
        const double om = 1.0 - lrho * lrho;
        const double star = 4.5 * kcap * L.beta * (om * om) * ds * proj;
        const bool right = cita == PI / 2.0;
        double fv[3];
        for (int k = 0; k < 3; ++k) fv[k] = star * (right ? nw2[k] : ntan[k]) * L.sigma;
        double tt[3];
        cross3(xr, fv, tt);
        for (int k = 0; k < 3; ++k) {
            acc[k] += fv[k];
            acc[3 + k] += tt[k];
        }
    }
    block_sum<7>(acc, Pt.part_lag + ((size_t)p * LAG_BLOCKS + blockIdx.y) * 7);
}

// Phases of the particles' movement step (k_part_sums, k_move).
enum { MV_MOM = 1, MV_CAP = 2, MV_REP = 4, MV_MOVE = 8, MV_ALL = 15 };

// ------------------------------------------------- per-particle sums
// Momentum-exchange (MV_MOM) and capillary (MV_CAP) force and torque of
// particle blockIdx.x: its slices' partial results staged in LDS by the whole
// workgroup, then each component summed over the slices in slice order by
// one thread (the order a serial sum takes, so every run gives the same
// bits).  MV_REP: the soft wall and pair repulsion (S:1602-1668), the pair
// terms of every other particle formed by the workgroup's threads at once
// (JC at a time, in LDS), then added in index order by one thread, as a
// serial loop over the pairs adds them.
static __global__ __launch_bounds__(PB) void k_part_sums(Lat L, Ext E, Part P, int phases)
{
    __shared__ double mom[LSLICES * 6], cap[LAG_BLOCKS * 7];
    constexpr int JC = 4 * PB;
    __shared__ double tj[JC][3];
    __shared__ unsigned char hit[JC];
    const int p = blockIdx.x;
    if (phases & MV_REP) {
        const int np = P.npar, n3[3] = {L.nx, L.ny, L.nz};
        const double range = 3.0, order = 0.01;
        const double R = P.rad[p];
        double cp[3], fr[3] = {0, 0, 0};
        for (int a = 0; a < 3; ++a) cp[a] = P.center[3 * p + a];
        if (threadIdx.x == 0)
            // walls (S:1606-1650): high faces push with +, low faces with -
            for (int a = 0; a < 3; ++a) {
                if (P.rep_on[2 * a]) {
                    const double d = cp[a] - P.rep_plane[2 * a] - R;
                    if (d < range) fr[a] -= 2.0 * (P.fg / order) * ((d - range) * (d - range)) / (range * range);
                }
                if (P.rep_on[2 * a + 1]) {
                    const double d = P.rep_plane[2 * a + 1] - cp[a] - R;
                    if (d < range) fr[a] += 2.0 * (P.fg / order) * ((d - range) * (d - range)) / (range * range);
                }
            }
        // pairs (S:1655-1668), minimum image on periodic axes
        for (int j0 = 0; j0 < np; j0 += JC) {
            const int jn = min(JC, np - j0);
            for (int t = threadIdx.x; t < jn; t += PB) {
                const int j = j0 + t;
                bool h = false;
                if (j != p) {
                    double dc[3];
                    for (int a = 0; a < 3; ++a) {
                        dc[a] = cp[a] - P.center[3 * j + a];
                        if (!(E.wall[a][0] || E.wall[a][1])) {
                            if (dc[a] > 0.5 * n3[a]) dc[a] -= n3[a];
                            if (dc[a] < -0.5 * n3[a]) dc[a] += n3[a];
                        }
                    }
                    const double d = sqrt(dc[0] * dc[0] + dc[1] * dc[1] + dc[2] * dc[2]) - (R + P.rad[j]);
                    if (d < range) {
                        h = true;
                        for (int a = 0; a < 3; ++a)
                            tj[t][a] = (P.fg / order) * ((d - range) * (d - range)) / (range * range) * dc[a] / d;
                    }
                }
                hit[t] = h;
            }
            __syncthreads();
            if (threadIdx.x == 0)
                for (int t = 0; t < jn; ++t)
                    if (hit[t])
                        for (int a = 0; a < 3; ++a) fr[a] = fr[a] - tj[t][a];
            __syncthreads();
        }
        if (threadIdx.x == 0)
            for (int a = 0; a < 3; ++a) P.frep[3 * p + a] = fr[a];
    }
    if (phases & MV_MOM)
        for (int t = threadIdx.x; t < LSLICES * 6; t += PB) mom[t] = P.part_mom[(size_t)p * LSLICES * 6 + t];
    if (phases & MV_CAP)
        for (int t = threadIdx.x; t < LAG_BLOCKS * 7; t += PB) cap[t] = P.part_lag[(size_t)p * LAG_BLOCKS * 7 + t];
    __syncthreads();
    const int k = threadIdx.x;
    if (k < 6 && (phases & MV_MOM)) {
        double m = 0;
        for (int s = 0; s < LSLICES; ++s) m += mom[s * 6 + k];
        if (k < 3)
            P.fp[3 * p + k] = m;
        else
            P.tp[3 * p + k - 3] = m;
    } else if (k >= 64 && k < 64 + 7 && (phases & MV_CAP)) { // a wave of its own
        const int kc = k - 64;
        double c = 0;
        for (int s = 0; s < LAG_BLOCKS; ++s) c += cap[s * 7 + kc];
        if (kc < 3)
            P.ten[3 * p + kc] = c;
        else if (kc < 6)
            P.tent[3 * p + kc - 3] = c;
        else
            P.shb[p] = c;
    }
}

// ------------------------------------------------------------ Newton-Euler
// One workgroup (MV_MOVE, after k_part_sums has finished the forces): the
// symmetry-plane rules and the explicit update of velocity, angular velocity
// and centre (S:1272-1346).
static __global__ __launch_bounds__(PB) void k_move(Lat L, Ext E, Part P, int phases)
{
    const int np = P.npar;
    if (!(phases & MV_MOVE)) return;
    if (P.sym_rules) {
        // the reference's symmetry-plane rules (S:1272-1335): a particle that
        // started on a mirror plane feels its image, so the in-plane force
        // components double and the normal ones vanish; particle par/3+1
        // (1-based) averages its x and z components
        const double xl = L.nx, yl = L.ny, zl = L.nz;
        for (int p = threadIdx.x; p < np; p += PB) {
            double *F = P.fp + 3 * p, *T = P.tp + 3 * p, *Fr = P.frep + 3 * p;
            double *te = P.ten + 3 * p, *tt = P.tent + 3 * p;
            const double *c0 = P.center0 + 3 * p;
            if (fabs(c0[0] - (xl - 0.5)) < (double)0.1f) {
                F[0] = 0.0; F[1] = 2.0 * F[1]; F[2] = 2.0 * F[2];
                T[0] = 2.0 * T[0]; T[1] = 0.0; T[2] = 0.0;
                Fr[0] = 0.0;
                te[0] = 0.0; te[1] = te[1] * 2.0; te[2] = te[2] * 2.0;
                tt[0] = 2.0 * tt[0]; tt[1] = 0.0; tt[2] = 0.0;
            }
            if (fabs(c0[1] - (yl - 0.5)) < (double)0.1f || fabs(c0[1] - 1.5) < (double)0.1f) {
                F[0] = 2.0 * F[0]; F[1] = 0.0; F[2] = 2.0 * F[2];
                T[0] = 0.0; T[1] = 2.0 * T[1]; T[2] = 0.0;
                Fr[1] = 0.0;
                te[0] = 2.0 * te[0]; te[1] = 0.0; te[2] = 2.0 * te[2];
                tt[0] = 0.0; tt[1] = 0.0; tt[2] = 0.0;
            }
            if (fabs(c0[2] - (zl - 0.5)) < (double)0.1f) {
                F[0] = 2.0 * F[0]; F[1] = 2.0 * F[1]; F[2] = 0.0;
                T[0] = 0.0; T[1] = 0.0; T[2] = 2.0 * T[2];
                Fr[2] = 0.0;
                te[0] = te[0] * 2.0; te[1] = te[1] * 2.0; te[2] = 0.0;
                tt[0] = 0.0; tt[1] = 0.0; tt[2] = 2.0 * tt[2];
            }
            if (p == np / 3) {
                double t = (te[0] + te[2]) / 2.0;
                te[0] = t;
                te[2] = t;
                tt[1] = 0.0;
                t = (F[0] + F[2]) / 2.0;
                F[0] = t;
                F[2] = t;
                T[1] = 0.0;
                t = (Fr[0] + Fr[2]) / 2.0;
                Fr[0] = t;
                Fr[2] = t;
            }
        }
        __syncthreads();
    }
    for (int p = threadIdx.x; p < np; p += PB) {
        for (int k = 0; k < 3; ++k) {
            const double ftot = P.fp[3 * p + k] + P.frep[3 * p + k] + P.ten[3 * p + k];
            P.up[3 * p + k] = P.up[3 * p + k] + ftot / P.mass[p];
            P.rp[3 * p + k] = P.rp[3 * p + k] + (P.tp[3 * p + k] + P.tent[3 * p + k]) / P.inertia[p];
            P.center_prev[3 * p + k] = P.center_old[3 * p + k];
            P.center_old[3 * p + k] = P.center[3 * p + k];
            P.center[3 * p + k] = P.center[3 * p + k] + P.up[3 * p + k] + 0.5 * ftot / P.mass[p];
        }
    }
}

// --------------------------------------------- ghost-layer phi and derivatives
// phi of the wall and mirror ghost layers = phi of the adjacent layer
// (S:806-807, S:1443-1465).  The reference copies whole planes face by face
// (walls z then x, then mirrors x, y, z), each copy reading planes the
// earlier ones already refreshed, so a ghost node on an edge or corner ends
// up with phi of the nearest live node; the source index clamps every ghost
// coordinate into the live range, which gives that result without ordering
// the six faces.
// With mirror_only, only the mirror faces are copied, clamping mirror axes
// only: the mirror fills that follow refilling and the ghost phi of live
// solid nodes (S:1397-1416, S:493-495).
template <class P> __global__ __launch_bounds__(256) void k_wall_phi(Lat L, Ext E, P *phi, int mirror_only = 0)
{
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    const int axis = blockIdx.y / 2, side = blockIdx.y % 2;
    if (mirror_only) {
        size_t d, s;
        unsigned A;
        if (mirror_pair(E, L, d, s, A)) phi[d] = phi[s];
        return;
    }
    if (!E.wall[axis][side] && !E.mirror[axis][side]) return;
    const int n3[3] = {L.nx, L.ny, L.nz};
    const int u = (axis + 1) % 3, v = (axis + 2) % 3;
    if (t >= (unsigned)(n3[u] * n3[v])) return;
    int cd[3], cs[3];
    cd[u] = (int)(t / n3[v]);
    cd[v] = (int)(t % n3[v]);
    cd[axis] = side ? n3[axis] - 1 : 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        cs[a] = cd[a];
        if ((E.wall[a][0] || E.mirror[a][0]) && cs[a] < 1) cs[a] = 1;
        if ((E.wall[a][1] || E.mirror[a][1]) && cs[a] > n3[a] - 2) cs[a] = n3[a] - 2;
    }
    phi[((size_t)cd[0] * L.ny + cd[1]) * L.nz + cd[2]] = phi[((size_t)cs[0] * L.ny + cs[1]) * L.nz + cs[2]];
}

// Derivatives of the current phi with the masks in force (S:498-587): zero at
// solid and ghost nodes, wetting at live fluid interface nodes.  Stores the
// NEAR band for the capillary kernel and, with out != null, every node as
// FP64 (gradx, grady, gradz, n_x, n_y, n_z) plus the flag.
template <class P>
__global__ __launch_bounds__(256) void k_derivs_ext(Lat L, Ext E, const P *__restrict__ phi, double *__restrict__ out,
                                                    int *__restrict__ flag_out)
{
    const int z = blockIdx.x * TZ + threadIdx.x;
    const int y = blockIdx.y * TY + threadIdx.y;
    const int x = blockIdx.z;
    if (z >= L.nz || y >= L.ny) return;
    unsigned xo[3], yo[3], zo[3];
    node_offsets(L, x, y, z, xo, yo, zo);
    const unsigned n = xo[1] + yo[1] + zo[1];
    const size_t N = (size_t)L.N;
    double g[3] = {0, 0, 0}, nv[3] = {0, 0, 0};
    bool fl = false;
    const unsigned nt = E.ntype ? E.ntype[n] : 0u;
    // a mirror ghost holds its image's gradient and normal (wetting
    // included), reflected (S:589-594); wall ghosts hold zeros
    const bool wg = is_wall_ghost(E, L, x, y, z);
    const unsigned A = wg ? 0u : mirror_axes(E, L, x, y, z);
    if (!wg && (A || !is_ghost(E, L, x, y, z))) {
        const int xi = (A & 1u) ? (x == 0 ? 1 : x - 1) : x, yi = (A & 2u) ? (y == 0 ? 1 : y - 1) : y,
                  zi = (A & 4u) ? (z == 0 ? 1 : z - 1) : z;
        unsigned ax[3], ay[3], az[3];
        node_offsets(L, xi, yi, zi, ax, ay, az);
        const unsigned ni = ax[1] + ay[1] + az[1];
        const unsigned nti = E.ntype ? E.ntype[ni] : 0u;
        if (!(nti & NT_SOLID)) {
            grad19<double>([&](int dx, int dy, int dz) { return double(phi[ax[1 + dx] + ay[1 + dy] + az[1 + dz]]); },
                           g[0], g[1], g[2]);
            const double fnorm = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
            if (fnorm > EVAP_LIM) {
                fl = A == 0;
                for (int k = 0; k < 3; ++k) nv[k] = g[k] / fnorm;
                if (nti & NT_POUT) {
                    const int pp = E.pout[ni] - 1;
                    wetting(L, E, xi + 1.0, yi + 1.0, zi + 1.0, E.center + 3 * pp, E.cita[pp], nv);
                    for (int k = 0; k < 3; ++k) g[k] = fnorm * nv[k];
                }
            }
            for (int k = 0; k < 3; ++k)
                if (A & (1u << k)) {
                    g[k] = -g[k];
                    nv[k] = -nv[k];
                }
        }
    }
    if ((nt & NT_NEAR) && E.der)
        for (int k = 0; k < 3; ++k) {
            E.der[k * N + n] = g[k];
            E.der[(3 + k) * N + n] = nv[k];
        }
    if (out) {
        for (int k = 0; k < 3; ++k) {
            out[k * N + n] = g[k];
            out[(3 + k) * N + n] = nv[k];
        }
        flag_out[n] = fl ? 1 : 0;
    }
}

} // namespace cgl
