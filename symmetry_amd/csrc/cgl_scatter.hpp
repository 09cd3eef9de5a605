// cgl_scatter.hpp -- the periodic FP32/FP16 factorised step with the next
// phase field summed in scatter form.
//
// The gather form computes phi of every node from its 19 upstream
// populations (the moments pass, calmacros S:770-831), which reads the whole
// state a second time per step.  Summed from the other side: right after its
// collision and recolouring (S:599-732) a node x knows everything it hands
// downstream.  f_i(x) goes to y = x + c_i, and its red part is
// r_i = A f_i + w_i G.c_i (S:713-726).  So the step kernel adds, for every
// downstream node y,
//     sR(y) += r_i(x),     sB(y) += f_i(x) - r_i(x),
// which are rho_R(y) and rho_B(y) of the next step (S:792-799), and
// phi(y) = (sR - sB) / (sR + sB).  The collision then takes
// rho_R = rho (1 + phi) / 2 with rho from its gathered populations, so the
// recolour scalars are needed only to rebuild red and blue populations for a
// download; they are written by the last step of a run.
//
// Sums are fixed point: each term is rounded to a multiple of rho0 * 2^-29
// and added as a 32-bit integer.  Integer sums are exact, so a node's sums do
// not depend on the order or grouping of their 19 terms, and results are the
// same for every workgroup box decomposition (x chunk).  Each term's
// minority colour is formed directly and the majority colour is the integer
// complement of the total, so a nearly pure node's small colour density is
// not swamped by the rounding of the large one; pure phases sum to an exact
// zero (A = 0 or 1, G = 0), so the EVAP_LIM rules of Recolor (S:720-725) see
// exact zeros there.  Densities up to 4 rho0 fit.
//
// Summation: a workgroup owns a box of TZ (z) x 8 (y) x xc (x) nodes and marches
// x.  A thread keeps its own node's sums for the destination planes x-1, x,
// x+1 in registers; z neighbours trade terms by lane shuffles, y neighbours
// through LDS.  Terms that leave the box go to per-node partial arrays, one
// per side class (x face low/high, y face, z face, x-y edge low/high, x-z edge
// low/high, y-z edge).  Every entry that a node's position calls for is written
// by exactly one workgroup each step, and the finishing pass adds them.
#pragma once

#include "cgl_kernels.hpp"
#include "cgl_particles.hpp"

namespace cgl {


// Arrays of N entries.  The first five kinds are indexed by node; the sums
// across a z face (fz, exzlo, exzhi, eyz) exist only for the first and last
// node of each 64-node z tile and are packed two per (x, y, z tile), at
// zslot(): a z-face read or write then touches 16 bytes per tile row instead
// of a cache line at each end of it.
struct ScBufs {
    int2 *own;           // sums from the node's own box
    int2 *fxlo, *fxhi;   // from the box on the node's low / high x side
    int2 *fy, *fz;       // from the box across the node's y / z face
    int2 *exylo, *exyhi; // from the box across an x-y edge, low / high x side
    int2 *exzlo, *exzhi; // from the box across an x-z edge, low / high x side
    int2 *eyz;           // from the box across a y-z edge
    int *range;          // set to 1 by a finishing pass that finds a colour sum wrapped (sticky)
};
constexpr int SC_NBUF = 10;

// Entries per (x, y) column of the packed z-face arrays: two per z tile (the
// tile's first node, side 0, and its last, side 1; tiles have two nodes or
// more), nx * ny * zcol_len <= N.
__device__ __forceinline__ unsigned zcol_len(const Lat &L) { return 2u * (unsigned)((L.nz + TZ - 1) / TZ); }
__device__ __forceinline__ unsigned zslot(const Lat &L, unsigned x, unsigned y, unsigned tile, unsigned side)
{
    return ((x * (unsigned)L.ny + y) * zcol_len(L)) + 2u * tile + side;
}

// Sums wrap modulo 2^32 (unsigned arithmetic: defined behaviour).  A colour
// density below 4 rho0 never wraps; one that reaches it lands below -rho0/2,
// where no sum of rounded terms of a real density can be, and sums_wrapped
// catches it.
__device__ __forceinline__ void add2(int2 &a, int2 b)
{
    a.x = (int)((unsigned)a.x + (unsigned)b.x);
    a.y = (int)((unsigned)a.y + (unsigned)b.y);
}
__device__ __forceinline__ bool sums_wrapped(int2 s) { return min(s.x, s.y) < -(1 << 28); }

// With particles the band kernel and the links add their terms to the
// destinations' own sums by one 64-bit atomic addition each, of
// (sB << 32) + sR with sR sign-extended.  That is exact for both halves when
// the red half of the final 64-bit value lies in [0, 2^32): the tile kernel
// writes every own red sum biased by SC_BIAS = 2^30 (a red partial or total
// lies in [-2^28, 2^31)), and the finishing pass removes the bias.
constexpr int SC_BIAS = 1 << 30;
__device__ __forceinline__ void atomic_add2(int2 *a, int2 c)
{
    const unsigned long long v =
        ((unsigned long long)(unsigned)c.y << 32) + (unsigned long long)(long long)c.x;
    atomicAdd((unsigned long long *)a, v);
}
__device__ __forceinline__ int2 shfl2(int2 v, int src) { return make_int2(__shfl(v.x, src), __shfl(v.y, src)); }

// ONE (all 19 direction arrays of a state within 4 GiB): the state as a raw
// buffer resource, so that a load or store takes the node's in-plane offset
// in a vector register and the plane and direction in a scalar one (one
// vector add per distinct in-plane offset instead of one per direction).
// 0x00020000: the gfx9 family's raw-buffer resource word.
template <class S> __device__ __forceinline__ __amdgpu_buffer_rsrc_t state_rsrc(const S *p, size_t n)
{
    return __builtin_amdgcn_make_buffer_rsrc((void *)p, 0, (int)(unsigned)(n * sizeof(S)), 0x00020000);
}
template <class S> __device__ __forceinline__ float buf_ld(__amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so)
{
    if constexpr (sizeof(S) == 4)
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, vo, so, 0));
    else
        return float(__builtin_bit_cast(_Float16, __builtin_amdgcn_raw_buffer_load_b16(r, vo, so, 0)));
}
// FP32 values streamed past the caches (non-temporal, the `nt` policy bit)
template <class S> __device__ __forceinline__ void buf_st(__amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so, float v)
{
    if constexpr (sizeof(S) == 4)
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, vo, so, 2);
    else
        __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, (_Float16)v), r, vo, so, 0);
}

// The phase-field buffer entry (signed minority fraction, minority_q) of a
// node from its sums: the minority sum over the total, IEEE division (a
// one-sided rounding of the colour split is amplified by the dynamics).  Pure
// phases give exactly -0 (red) and +0 (blue), phi = 1 and -1.
__device__ __forceinline__ float q_of_sums(int2 s)
{
    const int r = s.x > 0 ? s.x : 0, b = s.y > 0 ? s.y : 0;
    if (r == 0) return 0.0f;
    if (b == 0) return -0.0f;
    const float tot = float((unsigned)r + (unsigned)b); // exact below 2^32
    return b <= r ? -(float(b) / tot) : float(r) / tot;
}

// Shapes the scatter step takes with tiles of ky rows: every workgroup box
// has at least two rows and two lanes, so that its low and high faces are
// different threads.
__host__ __device__ inline bool sc_shape_ok(int ny, int nz, int ky = KTY<float>)
{
    return ny >= 2 && nz >= 2 && ny % ky != 1 && nz % TZ != 1;
}

// The box (z-tile bx, y-tile by, x-chunk bz) of one workgroup.  A tile at
// the high end of y or z may be partial: its last row / lane is the box face,
// and the threads beyond it take part in the shuffles and barriers only.
// ONE: all 19 direction arrays of a state lie within 4 GiB of its first, so
// every load and store takes the one base and a 32-bit offset (no per-direction
// 64-bit bases held in scalar registers).
// KY: rows of the tile (64 x KY threads).
// EXT (wall faces and particles, no mirror faces): the ghost layers and the
// particles' band (NEAR) are left to k_collide_sc_band (their terms are zero
// here); the normals are zero at wall ghosts and solid nodes and the wetted
// ones at the particles' link nodes; a node next to a wall adds the terms
// that leave through it to its own sums (half-way bounce-back returns them
// to the node, S:1557-1590) and takes no curvature force (S:624).
template <class S, bool ONE = false, int KY = KTY<float>, bool EXT = false>
__device__ __forceinline__ void collide_sc_tile(const Lat &L, const S *__restrict__ fin, S *__restrict__ fout,
                                                float4 *__restrict__ sout, const float *__restrict__ phi,
                                                const ScBufs &B, int xc, unsigned bx, unsigned by, unsigned bz,
                                                const Ext &E = Ext{})
{
    using R = float;
    __shared__ K2Smem<R, float, KY> sm;
    __shared__ int2 yx[2][3][KY][TZ]; // terms for the row above [0] / below [1], per x offset
    __shared__ int2 zdx[2][2][KY];    // edge-lane terms for the next row: [above/below][lane 0 / last][row]
    const int tz = threadIdx.x, ty = threadIdx.y, tid = ty * TZ + tz;
    const int z0 = bx * TZ, y0 = by * KY, x0 = bz * xc;
    const int x1 = min(x0 + xc, L.nx);
    const int zlast = min(TZ, L.nz - z0) - 1, ylast = min(KY, L.ny - y0) - 1; // last lane / row in the box
    const bool active = tz <= zlast && ty <= ylast;
    const int z = z0 + tz, y = y0 + ty;
    const unsigned sx = (unsigned)L.ny * (unsigned)L.nz;
    const size_t N = (size_t)L.N;
    const R rho0 = R(L.rho0), qs = SC_ONE / R(L.rho0);
    // (used with ONE only)
    const __amdgpu_buffer_rsrc_t rin = state_rsrc(fin, (size_t)Q * N);
    const __amdgpu_buffer_rsrc_t rout = state_rsrc((const S *)fout, (size_t)Q * N);

    // phi ring: each thread loads at most NE elements of the (KY+4) x (TZ+4)
    // plane region (the first always exists); their in-plane offsets are the
    // same for every plane
    constexpr int NPHI = (KY + 4) * (TZ + 4);
    constexpr int NE = (NPHI + TZ * KY - 1) / (TZ * KY);
    int pj[NE];
    unsigned po[NE];
#pragma unroll
    for (int k = 0; k < NE; ++k) {
        const int e = tid + k * TZ * KY;
        const int jy = e / (TZ + 4), jz = e - jy * (TZ + 4);
        pj[k] = e < NPHI ? e : -1;
        po[k] = e < NPHI ? (unsigned)wrapm(y0 - 2 + jy, L.ny) * (unsigned)L.nz + (unsigned)wrapm(z0 - 2 + jz, L.nz)
                         : 0u;
    }
    auto load_phi = [&](int p) {
        const unsigned gx = (unsigned)wrapm(p, L.nx) * sx;
        float *dst = &sm.phi[slot5(p)][0][0];
#pragma unroll
        for (int k = 0; k < NE; ++k)
            if (k == 0 || pj[k] >= 0) dst[pj[k]] = phi_of_q(at(phi, gx + po[k]));
    };
    // the ring's next plane is fetched one plane ahead into registers, so its
    // load latency overlaps a plane of work
    float pf[NE];
#pragma unroll
    for (int k = 0; k < NE; ++k) pf[k] = 0.f;
    auto prefetch_phi = [&](int p) {
        const unsigned gx = (unsigned)wrapm(p, L.nx) * sx;
#pragma unroll
        for (int k = 0; k < NE; ++k)
            if (k == 0 || pj[k] >= 0) pf[k] = at(phi, gx + po[k]);
    };
    auto store_phi = [&](int p) {
        float *dst = &sm.phi[slot5(p)][0][0];
#pragma unroll
        for (int k = 0; k < NE; ++k)
            if (k == 0 || pj[k] >= 0) dst[pj[k]] = phi_of_q(pf[k]);
    };
    // unit normals of plane p on the (KY+2) x (TZ+2) halo region (S:498-532).
    // Each thread computes its own node first (keeping that node's gradient)
    // and at most one element of the 1-node halo ring.
    constexpr int NHALO = 2 * (TZ + 2) + 2 * KY;
    static_assert(NHALO <= TZ * KY, "one halo element per thread");
    int hy = -1, hz = 0;
    if (tid < NHALO) {
        const int h = tid;
        if (h < TZ + 2) { hy = 0; hz = h; }
        else if (h < 2 * (TZ + 2)) { hy = KY + 1; hz = h - (TZ + 2); }
        else if (h < 2 * (TZ + 2) + KY) { hy = h - 2 * (TZ + 2) + 1; hz = 0; }
        else { hy = h - 2 * (TZ + 2) - KY + 1; hz = TZ + 1; }
    }
    // EXT: in-plane coordinates and offsets of this thread's node and of its
    // halo element, and their node types for the next plane of normals
    // (loaded a plane early, so that no load holds up a plane's start)
    // (bits of eflags: 0 and 1 whether the node in planes x and x + 1 is in
    // the band, 2 and 3 whether this node's / the halo element's column
    // lies in a y or z wall ghost layer, 4 whether the column is clear of
    // the y and z walls' force guard, 13-31 the node's bounce-back
    // directions off y and z walls (bits 0-18 of wall_bb_mask))
    const unsigned oin = (unsigned)wrapm(y, L.ny) * (unsigned)L.nz + (unsigned)wrapm(z, L.nz);
    const unsigned hin = (unsigned)wrapm(y0 - 1 + hy, L.ny) * (unsigned)L.nz + (unsigned)wrapm(z0 - 1 + hz, L.nz);
    unsigned eflags = 0;
    if constexpr (EXT) {
        auto wall_yz = [&](int py, int pz) {
            return (E.wall[1][0] && py == 0) || (E.wall[1][1] && py == L.ny - 1) || (E.wall[2][0] && pz == 0) ||
                   (E.wall[2][1] && pz == L.nz - 1);
        };
        const int yw0 = wrapm(y, L.ny), zw0 = wrapm(z, L.nz);
        if (wall_yz(yw0, zw0)) eflags |= 4u;
        if (wall_yz(wrapm(y0 - 1 + hy, L.ny), wrapm(z0 - 1 + hz, L.nz))) eflags |= 8u;
        if (!((E.wall[1][0] && yw0 < 2) || (E.wall[1][1] && yw0 > L.ny - 3) || (E.wall[2][0] && zw0 < 2) ||
              (E.wall[2][1] && zw0 > L.nz - 3)))
            eflags |= 16u;
        eflags |= (wall_bb_axis(E, 1, yw0, L.ny) | wall_bb_axis(E, 2, zw0, L.nz)) << 13;
    }
    unsigned nto = 0, nth = 0;
    auto load_nt = [&](int p) {
        if constexpr (EXT) {
            if (E.ntype) {
                const unsigned gxp = (unsigned)wrapm(p, L.nx) * sx;
                nto = at(E.ntype, gxp + oin);
                if (hy >= 0) nth = at(E.ntype, gxp + hin);
            }
        }
    };
    auto normal_at = [&](int p, int jy, int jz, R &gx, R &gy, R &gz, bool wyz, unsigned pin, unsigned nt) {
        const int s = slot3(p);
        grad19<R>([&](int dx, int dy, int dz) { return sm.phi[slot5(p + dx)][jy + 1 + dy][jz + 1 + dz]; }, gx, gy,
                  gz);
        const R g2 = gx * gx + gy * gy + gz * gz;
        R nx = 0, ny = 0, nz = 0;
        if (g2 > R(EVAP_LIM * EVAP_LIM)) {
            const R inv = __builtin_amdgcn_rsqf(g2);
            nx = gx * inv;
            ny = gy * inv;
            nz = gz * inv;
        }
        if constexpr (EXT) {
            // (S:498-532, S:536-594) zero at wall ghosts and solid nodes, the
            // wetted normal that k_wet_normals stored at link nodes
            const int px = wrapm(p, L.nx);
            const bool zero =
                wyz || (E.wall[0][0] && px == 0) || (E.wall[0][1] && px == L.nx - 1) || (nt & NT_SOLID);
            if (!zero && (nt & NT_POUT) && g2 > R(EVAP_LIM * EVAP_LIM)) {
                const unsigned m = (unsigned)px * sx + pin;
                nx = R(at(E.der, 3 * N + m));
                ny = R(at(E.der, 4 * N + m));
                nz = R(at(E.der, 5 * N + m));
            }
            if (zero) nx = ny = nz = 0;
        }
        sm.n[0][s][jy][jz] = nx;
        sm.n[1][s][jy][jz] = ny;
        sm.n[2][s][jy][jz] = nz;
    };
    auto normals = [&](int p, R &gx, R &gy, R &gz) {
        normal_at(p, ty + 1, tz + 1, gx, gy, gz, (eflags & 4u) != 0, oin, nto);
        if (hy >= 0) {
            R a, b, c;
            normal_at(p, hy, hz, a, b, c, (eflags & 8u) != 0, hin, nth);
        }
    };

    // where this thread's box-leaving terms go (boxes have at least two rows
    // and two lanes, so the low and high faces are different threads)
    const bool ylo = ty == 0, yhi = ty == ylast, yedge = active && (ylo || yhi);
    // lane 0 holds the terms for z0 + zlast + 1, the last lane those for z0 - 1
    const bool zl = tz == 0, zh = tz == zlast, zedge = active && (zl || zh);
    const int left = zl ? zlast : tz - 1, right = zh ? 0 : (tz + 1) & (TZ - 1);
    const int zs = zl ? 0 : 1;
    // packed z-face slot of the destination: lane 0 feeds the first node of
    // the next z tile, the last lane the last node of the previous one
    const unsigned nzt = (unsigned)(L.nz + TZ - 1) / TZ;
    const unsigned zc2 = zl ? 2u * (bx + 1 == nzt ? 0u : bx + 1) : 2u * (bx == 0 ? nzt - 1 : bx - 1) + 1u;
    const unsigned yw = (unsigned)wrapm(y, L.ny), yo = (unsigned)(ylo ? wrapm(y0 - 1, L.ny) : wrapm(y0 + ylast + 1, L.ny));
    // in-plane offsets: this node, the node outside the y face next to it, and
    // the packed z-face slots of this row and of that outside row
    const unsigned nself = yw * (unsigned)L.nz + (unsigned)wrapm(z, L.nz);
    const unsigned nrow = yo * (unsigned)L.nz + (unsigned)wrapm(z, L.nz);
    const unsigned cself = yw * zcol_len(L) + zc2, crow = yo * zcol_len(L) + zc2;
    const unsigned csx = (unsigned)L.ny * zcol_len(L); // packed z-face plane stride

    // sums for destination planes x-1 (m), x (c), x+1 (p): own node column,
    // the row outside the box next to this row, the column outside next to this lane
    const int2 Z2 = make_int2(0, 0);
    int2 am = Z2, ac = Z2, ap = Z2;
    int2 ym = Z2, yc = Z2, yp = Z2;
    int2 zm = Z2, zc = Z2, zp = Z2;

    for (int p = x0 - 2; p <= x0 + 2; ++p) load_phi(p);
    __syncthreads();
    R gx, gy, gz; // gradient of phi at the node, plane x
    load_nt(x0 - 1);
    {
        R a, b, c;
        normals(x0 - 1, a, b, c);
    }
    load_nt(x0);
    normals(x0, gx, gy, gz);
    if (nto & NT_NEAR) eflags |= 1u;
    load_nt(x0 + 1);

    for (int x = x0; x < x1; ++x) {
        __syncthreads(); // readers of the slots refilled below are done
        if (x > x0) store_phi(x + 2);
        prefetch_phi(x + 3);
        __syncthreads();
        R ngx, ngy, ngz;
        normals(x + 1, ngx, ngy, ngz);
        if constexpr (EXT)
            if (nto & NT_NEAR) eflags |= 2u;
        __syncthreads();

        // terms (rho_R, rho_B) for the downstream nodes, one direction per
        // destination offset (dx, dy, dz): g00 (dx,0,0), gzp/gzm (dx,0,+-1),
        // gyp/gym (dx,+-1,0), dpp/dpm/dmp/dmm (0,+-1,+-1); zero outside the box
        int2 g00[3] = {Z2, Z2, Z2}, gzp[3] = {Z2, Z2, Z2}, gzm[3] = {Z2, Z2, Z2}, gyp[3] = {Z2, Z2, Z2},
             gym[3] = {Z2, Z2, Z2}, dpp = Z2, dpm = Z2, dmp = Z2, dmm = Z2;
        bool live = active;
        unsigned wbb = 0; // wall bounce-back directions of the node
        if constexpr (EXT) {
            // (the x walls' parts are the same for the whole plane)
            if (live)
                live = !(eflags & 5u) && !(E.wall[0][0] && x == 0) && !(E.wall[0][1] && x == L.nx - 1);
            if (live) wbb = (eflags >> 13) | wall_bb_axis(E, 0, x, L.nx);
            eflags = (eflags & ~3u) | ((eflags >> 1) & 1u);
        }
        if (live) {
            // gradient, normal and flag at the node (S:498-532)
            const R g2 = gx * gx + gy * gy + gz * gz;
            const bool flag = g2 > R(EVAP_LIM * EVAP_LIM);
            bool force_on = flag;
            if constexpr (EXT)
                force_on = flag && (eflags & 16u) && !((E.wall[0][0] && x < 2) || (E.wall[0][1] && x > L.nx - 3));
            const R nx = sm.n[0][slot3(x)][ty + 1][tz + 1];
            const R ny = sm.n[1][slot3(x)][ty + 1][tz + 1];
            const R nz = sm.n[2][slot3(x)][ty + 1][tz + 1];
            // normal derivatives for the curvature (S:635-660)
            R d[3][3];
            if (force_on)
                curv19<R>([&](int k, int dx, int dy, int dz) { return sm.n[k][slot3(x + dx)][ty + 1 + dy][tz + 1 + dz]; },
                          d);

            // streaming: pull gather of the total populations (deviations from w rho0)
            unsigned xo[3], yo[3], zo[3];
            node_offsets(L, x, y, z, xo, yo, zo);
            const unsigned n = xo[1] + yo[1] + zo[1];
            R f[Q];
            R df = 0, jx = 0, jy = 0, jz = 0;
            // FP32 storage: the wave issues its 19 loads at raised priority, so
            // that the other waves of the SIMD do not delay them (+0.8 % MLUPS
            // at 256x256x512; neutral for FP16 storage)
            if constexpr (sizeof(S) == 4) __builtin_amdgcn_s_setprio(2);
#pragma unroll
            for (int i = 0; i < Q; ++i) {
                if constexpr (ONE) {
                    const unsigned vo = (yo[1 - CY[i]] + zo[1 - CZ[i]]) * (unsigned)sizeof(S);
                    const unsigned so = (xo[1 - CX[i]] + (unsigned)i * (unsigned)N) * (unsigned)sizeof(S);
                    f[i] = buf_ld<S>(rin, vo, so);
                } else {
                    const unsigned m = xo[1 - CX[i]] + yo[1 - CY[i]] + zo[1 - CZ[i]];
                    f[i] = StoreOf<S>::ld(grp(fin, i, N), grp_off(i, N, m));
                }
            }
            if constexpr (sizeof(S) == 4) __builtin_amdgcn_s_setprio(0);
#pragma unroll
            for (int i = 0; i < Q; ++i) {
                df += f[i];
                if (CX[i] == 1) jx += f[i];
                if (CX[i] == -1) jx -= f[i];
                if (CY[i] == 1) jy += f[i];
                if (CY[i] == -1) jy -= f[i];
                if (CZ[i] == 1) jz += f[i];
                if (CZ[i] == -1) jz -= f[i];
            }
            // moments (S:792-799): rho and j from the gathered populations, the
            // colour split from the scattered phase field
            const R phin = sm.phi[slot5(x)][ty + 2][tz + 2];
            Moments<R, true> mo;
            mo.rho = rho0 + df;
            mo.drho = df;
            const R irho = fdiv(R(1), mo.rho);
            mo.ux = jx * irho;
            mo.uy = jy * irho;
            mo.uz = jz * irho;
            mo.phi = phin;
            collide<R, true>(L, f, mo, phin, gx, gy, gz, nx, ny, nz, force_on, d);
            // the colour split (the collision needs only rho, u and phi): the
            // minority colour's fraction hm from the phase-field buffer (its
            // signed minority fraction), so that a nearly pure node keeps its
            // small colour density to FP32's relative precision; the majority
            // colour is the complement
            const float qn = at(phi, (unsigned)x * sx + nself);
            const bool redmaj = __builtin_signbitf(qn);
            const R hm = __builtin_fabsf(qn);
            const R smin = mo.rho * hm;
            mo.sr = redmaj ? mo.rho - smin : smin;
            mo.sb = redmaj ? smin : mo.rho - smin;
            const Recol<R> rc = recol<R, true>(L, mo, flag);
            // recolouring r_i = A f_i + g w_i n.c_i (S:713-726), with the
            // EVAP_LIM overrides; the minority colour's part of f_i is
            // Mf f_i + w_i (Mg n).c_i (red: A = hm, g; blue: 1 - A = hm, -g)
            R g = rc.gam;
            R Mf = hm;
            if (rc.evr) { g = 0; Mf = redmaj ? R(1) : R(0); }
            if (rc.evb) { g = 0; Mf = redmaj ? R(0) : R(1); }
            const R Gx = g * nx, Gy = g * ny, Gz = g * nz;
            const R Mg = redmaj ? -g : g;

            // the next state: post-collision totals, and the recolour scalars on
            // a run's last step; streamed past the caches
#pragma unroll
            for (int i = 0; i < Q; ++i) {
                if constexpr (ONE) {
                    buf_st<S>(rout, (yo[1] + zo[1]) * (unsigned)sizeof(S),
                              (xo[1] + (unsigned)i * (unsigned)N) * (unsigned)sizeof(S), float(f[i]));
                } else {
                    S *fp = grp(fout, i, N);
                    const unsigned k = grp_off(i, N, n);
                    if constexpr (sizeof(S) == 4)
                        __builtin_nontemporal_store(float(f[i]), (float *)&at(fp, k));
                    else
                        StoreOf<S>::st(fp, k, f[i]);
                }
            }
            if constexpr (EXT) {
                // the wall ghost slots this node reads at the next step: slot i of
                // the ghost x - c_i holds the node's own opposite population (the
                // half-way bounce-back, S:1557-1590), which the gather reads there
                if (wbb) {
                    // (offsets formed again here, behind an empty asm that keeps
                    // the compiler from reusing, and so holding, the gather's)
                    int yy = y, zz = z;
                    asm volatile("" : "+v"(yy), "+v"(zz));
                    unsigned gxo[3], gyo[3], gzo[3];
                    node_offsets(L, x, yy, zz, gxo, gyo, gzo);
#pragma unroll
                    for (int i = 1; i < Q; ++i) {
                        if (!((wbb >> i) & 1u)) continue;
                        const unsigned m = gxo[1 - CX[i]] + gyo[1 - CY[i]] + gzo[1 - CZ[i]];
                        StoreOf<S>::st(grp(fout, i, N), grp_off(i, N, m), f[OPC[i]]);
                    }
                }
            }
            if (sout) {
                typedef float f4v __attribute__((ext_vector_type(4)));
                Recol<R> ce = rc;
                ce.M = hm;
                ce.redmaj = redmaj;
                const f4v v4 = {encode_frac(ce), float(Gx), float(Gy), float(Gz)};
                __builtin_nontemporal_store(v4, (f4v *)&at(sout, n));
            }

            // the terms in fixed-point units, rounded to the nearest integer,
            // halves up (a one-sided rounding such as truncation would bias
            // every sum by up to 19 units): the total f_i = f~_i qs + w_i 2^29,
            // the minority colour's part Mf f_i + (w_i qs Mg n).c_i, small where
            // the node is nearly pure, and the majority colour's part as the
            // integer complement.  (At 2^29 units an FP32 value resolves only
            // 2-16 units, so a majority term formed in FP32 would leave a
            // minority complement of noise.)  Both colour parts take the same
            // FP32 total f_i, so the rounding of its constant w_i 2^29 splits in
            // proportion to the colour fractions and leaves phi unbiased.
            const R s1 = R(W[1]) * qs, s7 = R(W[7]) * qs;
            const R Mgx = Mg * nx, Mgy = Mg * ny, Mgz = Mg * nz;
            const R G1x = Mgx * s1, G1y = Mgy * s1, G1z = Mgz * s1;
            const R G7x = Mgx * s7, G7y = Mgy * s7, G7z = Mgz * s7;
#pragma unroll
            for (int i = 0; i < Q; ++i) {
                R v = f[i];
                if constexpr (sizeof(S) != 4) v = R(S(v)); // the value the next step reads
                const R fq = __builtin_fmaf(v, qs, R(W[i] * double(SC_ONE)));
                R cq = Mf * fq;
                if (i > 0 && i <= 6) cq = __builtin_fmaf(Mf, fq, dot_c<R>(i, G1x, G1y, G1z));
                if (i > 6) cq = __builtin_fmaf(Mf, fq, dot_c<R>(i, G7x, G7y, G7z));
                const int qf = round_half_up(fq), qm = round_half_up(cq);
                const int2 c = redmaj ? make_int2(qf - qm, qm) : make_int2(qm, qf - qm);
                const int k = CX[i] + 1;
                if (CY[i] == 0 && CZ[i] == 0) g00[k] = c;
                else if (CY[i] == 0 && CZ[i] == 1) gzp[k] = c;
                else if (CY[i] == 0) gzm[k] = c;
                else if (CZ[i] == 0 && CY[i] == 1) gyp[k] = c;
                else if (CZ[i] == 0) gym[k] = c;
                else if (CY[i] == 1 && CZ[i] == 1) dpp = c;
                else if (CY[i] == 1) dpm = c;
                else if (CZ[i] == 1) dmp = c;
                else dmm = c;
            }
            if constexpr (EXT) {
                if (wbb) {
                    // terms into a wall ghost return to the node itself
#pragma unroll
                    for (int i = 1; i < Q; ++i) {
                        if (!((wbb >> OPC[i]) & 1u)) continue;
                        const int k = CX[i] + 1;
                        int2 &t = (CY[i] == 0 && CZ[i] == 0) ? g00[k]
                                  : (CY[i] == 0 && CZ[i] == 1) ? gzp[k]
                                  : (CY[i] == 0)               ? gzm[k]
                                  : (CZ[i] == 0 && CY[i] == 1) ? gyp[k]
                                  : (CZ[i] == 0)               ? gym[k]
                                  : (CY[i] == 1 && CZ[i] == 1) ? dpp
                                  : (CY[i] == 1)               ? dpm
                                  : (CZ[i] == 1)               ? dmp
                                                               : dmm;
                        add2(g00[1], t);
                        t = Z2;
                    }
                }
            }
        }

        // z neighbours: lane shuffles (every lane takes part).  Lane 0 receives
        // the last lane's +z terms and the last lane lane 0's -z terms: those
        // leave the box.
        int2 own[3] = {g00[0], g00[1], g00[2]};
        int2 zo3[3] = {Z2, Z2, Z2}, zup = Z2, zdn = Z2;
        {
            int2 lz[3], rz[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                lz[k] = shfl2(gzp[k], left);
                rz[k] = shfl2(gzm[k], right);
            }
            const int2 lpp = shfl2(dpp, left), lmp = shfl2(dmp, left);
            const int2 rpm = shfl2(dpm, right), rmm = shfl2(dmm, right);
            if (!zl) {
#pragma unroll
                for (int k = 0; k < 3; ++k) add2(own[k], lz[k]);
                add2(gyp[1], lpp);
                add2(gym[1], lmp);
            } else {
#pragma unroll
                for (int k = 0; k < 3; ++k) zo3[k] = lz[k];
                zup = lpp;
                zdn = lmp;
            }
            if (!zh) {
#pragma unroll
                for (int k = 0; k < 3; ++k) add2(own[k], rz[k]);
                add2(gyp[1], rpm);
                add2(gym[1], rmm);
            } else {
#pragma unroll
                for (int k = 0; k < 3; ++k) zo3[k] = rz[k];
                zup = rpm;
                zdn = rmm;
            }
        }

        // y neighbours through LDS; y-z edge terms of the corner lanes go out
        // at once (they have dx = 0 only)
        if (active && !yhi) {
#pragma unroll
            for (int k = 0; k < 3; ++k) yx[0][k][ty][tz] = gyp[k];
        }
        if (active && !ylo) {
#pragma unroll
            for (int k = 0; k < 3; ++k) yx[1][k][ty][tz] = gym[k];
        }
        if (zedge) {
            if (!yhi) zdx[0][zs][ty] = zup;
            else at(B.eyz, (unsigned)x * csx + crow) = zup;
            if (!ylo) zdx[1][zs][ty] = zdn;
            else at(B.eyz, (unsigned)x * csx + crow) = zdn;
        }
        __syncthreads();
        if (active && !ylo) {
#pragma unroll
            for (int k = 0; k < 3; ++k) add2(own[k], yx[0][k][ty - 1][tz]);
        }
        if (active && !yhi) {
#pragma unroll
            for (int k = 0; k < 3; ++k) add2(own[k], yx[1][k][ty + 1][tz]);
        }
        if (zedge) {
            if (!ylo) add2(zo3[1], zdx[0][zs][ty - 1]);
            if (!yhi) add2(zo3[1], zdx[1][zs][ty + 1]);
        }

        // accumulate over source planes; destination plane x-1 is complete
        add2(am, own[0]);
        add2(ac, own[1]);
        add2(ap, own[2]);
        if (yedge) {
            add2(ym, ylo ? gym[0] : gyp[0]);
            add2(yc, ylo ? gym[1] : gyp[1]);
            add2(yp, ylo ? gym[2] : gyp[2]);
        }
        if (zedge) {
            add2(zm, zo3[0]);
            add2(zc, zo3[1]);
            add2(zp, zo3[2]);
        }
        const unsigned pm = (unsigned)wrapm(x - 1, L.nx) * sx;
        if (!active) {
        } else if (x > x0) {
            at(B.own, pm + nself) = EXT ? make_int2((int)((unsigned)am.x + (unsigned)SC_BIAS), am.y) : am;
            if (yedge) at(B.fy, pm + nrow) = ym;
            if (zedge) at(B.fz, (unsigned)wrapm(x - 1, L.nx) * csx + cself) = zm;
        } else { // plane x0 - 1 belongs to the box on the low x side
            at(B.fxhi, pm + nself) = am;
            if (yedge) at(B.exyhi, pm + nrow) = ym;
            if (zedge) at(B.exzhi, (unsigned)wrapm(x - 1, L.nx) * csx + cself) = zm;
        }
        load_nt(x + 2);
        gx = ngx;
        gy = ngy;
        gz = ngz;
        am = ac;
        ac = ap;
        ap = Z2;
        ym = yc;
        yc = yp;
        yp = Z2;
        zm = zc;
        zc = zp;
        zp = Z2;
    }
    // plane x1 - 1 (own box) and plane x1 (the box on the high x side)
    if (!active) return;
    const unsigned pl = (unsigned)wrapm(x1 - 1, L.nx) * sx, pn = (unsigned)wrapm(x1, L.nx) * sx;
    at(B.own, pl + nself) = EXT ? make_int2((int)((unsigned)am.x + (unsigned)SC_BIAS), am.y) : am;
    at(B.fxlo, pn + nself) = ac;
    if (yedge) {
        at(B.fy, pl + nrow) = ym;
        at(B.exylo, pn + nrow) = yc;
    }
    if (zedge) {
        at(B.fz, (unsigned)wrapm(x1 - 1, L.nx) * csx + cself) = zm;
        at(B.exzlo, (unsigned)wrapm(x1, L.nx) * csx + cself) = zc;
    }
}

template <class S, bool ONE = false, int KY = KTY<float>>
__global__ __launch_bounds__(TZ *KY, 4) void k_collide_sc(Lat L, const S *__restrict__ fin, S *__restrict__ fout,
                                                          float4 *__restrict__ sout, const float *__restrict__ phi,
                                                          ScBufs B, int xc)
{
    unsigned bx, by, bz;
    tile_of_block(L, bx, by, bz);
    collide_sc_tile<S, ONE, KY>(L, fin, fout, sout, phi, B, xc, bx, by, bz);
}

// The scatter step with wall faces and particles: every lane but the ghost
// layers and the particles' band.
template <class S, bool ONE>
__global__ __launch_bounds__(TZ *KTY<float>, 4) void k_collide_sc_ext(Lat L, const S *__restrict__ fin,
                                                                    S *__restrict__ fout, float4 *__restrict__ sout,
                                                                    const float *__restrict__ phi, ScBufs B, int xc,
                                                                    Ext E)
{
    unsigned bx, by, bz;
    tile_of_block(L, bx, by, bz);
    collide_sc_tile<S, ONE, KTY<float>, true>(L, fin, fout, sout, phi, B, xc, bx, by, bz, E);
}

// Partial-sum arrays of box `base` in a packed batch.
__device__ __forceinline__ ScBufs sc_offset(const ScBufs &B, unsigned long long base)
{
    return ScBufs{B.own + base,   B.fxlo + base,  B.fxhi + base,  B.fy + base,    B.fz + base,
                  B.exylo + base, B.exyhi + base, B.exzlo + base, B.exzhi + base, B.eyz + base, B.range};
}

// Every box of a batch in one launch; a workgroup finds its box by the
// first-tile table, and the box's arrays start at its packed offset.
template <class S>
__global__ __launch_bounds__(TZ *KTY<float>, 4) void k_collide_sc_batch(const BoxDev *__restrict__ BX, int nbox, int remap,
                                                                      int xc, const S *__restrict__ fin,
                                                                      S *__restrict__ fout,
                                                                      float4 *__restrict__ sout,
                                                                      const float *__restrict__ phi, ScBufs B)
{
    const unsigned t = remap ? xcd_order(blockIdx.x, gridDim.x) : blockIdx.x;
    const int b = find_box<true>(BX, nbox, t);
    const Lat L = BX[b].L;
    const unsigned long long base = BX[b].base;
    const unsigned lt = t - BX[b].t2, gz = BX[b].gz, gy = BX[b].gy2;
    collide_sc_tile<S>(L, fin + (size_t)Q * base, fout + (size_t)Q * base, sout ? sout + base : nullptr,
                       phi + base, sc_offset(B, base), xc, lt % gz, (lt / gz) % gy,
                       lt / (gz * gy));
}

// Finishing pass at node (x, y, z): phi(y) = (sR - sB) / (sR + sB) from the
// node's own-box sums and the partial sums its position in its workgroup box
// (x chunk xc, ky rows) calls for.
__device__ __forceinline__ void phi_sc_node(const Lat &L, const ScBufs &B, int xc, int ky, float *__restrict__ phi,
                                            int x, int y, int z, int bias = 0)
{
    const unsigned n = ((unsigned)x * (unsigned)L.ny + (unsigned)y) * (unsigned)L.nz + (unsigned)z;
    int2 s = at(B.own, n);
    s.x = (int)((unsigned)s.x - (unsigned)bias);
    const int cx0 = (x / xc) * xc;
    const bool xlo = x == cx0, xhi = x == min(cx0 + xc, L.nx) - 1;
    const int yr = y % ky, zr = z % TZ;
    const bool fy = yr == 0 || y == min(y - yr + ky, L.ny) - 1;
    const bool fz = zr == 0 || z == min(z - zr + TZ, L.nz) - 1;
    if (xlo) add2(s, at(B.fxlo, n));
    if (xhi) add2(s, at(B.fxhi, n));
    if (fy) add2(s, at(B.fy, n));
    if (fy && xlo) add2(s, at(B.exylo, n));
    if (fy && xhi) add2(s, at(B.exyhi, n));
    if (fz) {
        const unsigned c = zslot(L, (unsigned)x, (unsigned)y, (unsigned)(z / TZ), zr == 0 ? 0u : 1u);
        add2(s, at(B.fz, c));
        if (xlo) add2(s, at(B.exzlo, c));
        if (xhi) add2(s, at(B.exzhi, c));
        if (fy) add2(s, at(B.eyz, c));
    }
    if (sums_wrapped(s)) *B.range = 1;
    at(phi, n) = q_of_sums(s);
}

// bias: SC_BIAS where the own red sums carry it (walls and particles), else 0.
static __global__ __launch_bounds__(256) void k_phi_sc(Lat L, ScBufs B, int xc, int ky, float *__restrict__ phi,
                                                       int bias)
{
    const int z = blockIdx.x * TZ + threadIdx.x, y = blockIdx.y * TY + threadIdx.y, x = blockIdx.z;
    if (z >= L.nz || y >= L.ny) return;
    phi_sc_node(L, B, xc, ky, phi, x, y, z, bias);
}

// The finishing pass two nodes at a time, the pair (x, y, z), (x, y, z + 1)
// (z even, nz even): the pair's own sums in one 16-byte load (o, loaded by
// the caller), the x-face and y-face sums (which call for both nodes of the
// pair or for neither) likewise, the z-face sums per node, phi in one 8-byte
// store.  Integer sums are exact, so the order of the additions does not
// change phi.
__device__ __forceinline__ void phi_sc_pair(const Lat &L, const ScBufs &B, int xc, int ky, float *__restrict__ phi,
                                            int bias, int x, int y, int z, unsigned h, int4 o)
{
    int2 s[2] = {make_int2((int)((unsigned)o.x - (unsigned)bias), o.y),
                 make_int2((int)((unsigned)o.z - (unsigned)bias), o.w)};
    auto add_pair = [&](const int2 *a) {
        const int4 v = at((const int4 *)a, h);
        add2(s[0], make_int2(v.x, v.y));
        add2(s[1], make_int2(v.z, v.w));
    };
    const int cx0 = (x / xc) * xc;
    const bool xlo = x == cx0, xhi = x == min(cx0 + xc, L.nx) - 1;
    const int yr = y % ky;
    const bool fy = yr == 0 || y == min(y - yr + ky, L.ny) - 1;
    if (xlo) add_pair(B.fxlo);
    if (xhi) add_pair(B.fxhi);
    if (fy) add_pair(B.fy);
    if (fy && xlo) add_pair(B.exylo);
    if (fy && xhi) add_pair(B.exyhi);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int zk = z + k, zr = zk % TZ;
        if (zr == 0 || zk == min(zk - zr + TZ, L.nz) - 1) {
            const unsigned c = zslot(L, (unsigned)x, (unsigned)y, (unsigned)(zk / TZ), zr == 0 ? 0u : 1u);
            add2(s[k], at(B.fz, c));
            if (xlo) add2(s[k], at(B.exzlo, c));
            if (xhi) add2(s[k], at(B.exzhi, c));
            if (fy) add2(s[k], at(B.eyz, c));
        }
    }
    if (sums_wrapped(s[0]) || sums_wrapped(s[1])) *B.range = 1;
    at((float2 *)phi, h) = make_float2(q_of_sums(s[0]), q_of_sums(s[1]));
}

static __global__ __launch_bounds__(256) void k_phi_sc2(Lat L, ScBufs B, int xc, int ky, float *__restrict__ phi,
                                                        int bias)
{
    const int z = (blockIdx.x * TZ + threadIdx.x) * 2, y = blockIdx.y * TY + threadIdx.y, x = blockIdx.z;
    if (z >= L.nz || y >= L.ny) return;
    const unsigned h = (((unsigned)x * (unsigned)L.ny + (unsigned)y) * (unsigned)L.nz + (unsigned)z) >> 1;
    phi_sc_pair(L, B, xc, ky, phi, bias, x, y, z, h, at((const int4 *)B.own, h));
}

// The same, two x planes per thread (nx even): both planes' own sums are
// loaded before either pair is finished, two 16-byte loads in flight per
// thread.
static __global__ __launch_bounds__(256) void k_phi_sc4(Lat L, ScBufs B, int xc, int ky, float *__restrict__ phi,
                                                        int bias)
{
    const int z = (blockIdx.x * TZ + threadIdx.x) * 2, y = blockIdx.y * TY + threadIdx.y, x = blockIdx.z * 2;
    if (z >= L.nz || y >= L.ny) return;
    const unsigned h = (((unsigned)x * (unsigned)L.ny + (unsigned)y) * (unsigned)L.nz + (unsigned)z) >> 1;
    const unsigned h2 = h + (((unsigned)L.ny * (unsigned)L.nz) >> 1);
    const int4 o = at((const int4 *)B.own, h), o2 = at((const int4 *)B.own, h2);
    phi_sc_pair(L, B, xc, ky, phi, bias, x, y, z, h, o);
    phi_sc_pair(L, B, xc, ky, phi, bias, x + 1, y, z, h2, o2);
}

// The batch's finishing pass over the moments launch's tiles (z tile, y tile,
// chunk of xs planes) of every box.
static __global__ __launch_bounds__(256) void k_phi_sc_batch(const BoxDev *__restrict__ BX, int nbox, int xs, int xc,
                                                             ScBufs B, float *__restrict__ phi)
{
    const unsigned t = blockIdx.x;
    const int b = find_box<false>(BX, nbox, t);
    const Lat L = BX[b].L;
    const unsigned long long base = BX[b].base;
    const unsigned lt = t - BX[b].t1, gz = BX[b].gz, gy = BX[b].gy;
    const int z = (int)(lt % gz) * TZ + threadIdx.x, y = (int)((lt / gz) % gy) * TY + threadIdx.y;
    const int xb = (int)(lt / (gz * gy)) * xs;
    if (z >= L.nz || y >= L.ny) return;
    const ScBufs Bb = sc_offset(B, base);
    for (int x = xb; x < min(xb + xs, L.nx); ++x) phi_sc_node(L, Bb, xc, KTY<float>, phi + base, x, y, z);
}

// ------------------------------------------------- the particles' band
// With particles, the scatter step's tile kernel leaves the band (NEAR: the
// solid nodes and every node within R + 3 of a centre) to the kernels below:
// one thread per node of particle blockIdx.y's band box (R + 3), node
// blockIdx.x PB + threadIdx.x in box order.

// Whether box node (x, y, z) of particle blockIdx.y lies in no lower
// particle's band box, so that a node in several boxes adds its terms once.
// Every thread of the workgroup takes part (barriers): the lower particles'
// boxes that meet this particle's own pass through LDS, PB candidates at a
// time, their periodic corners wrapped into the domain, and only those are
// tested per node (the answer does not depend on their order in the list).
__device__ __forceinline__ bool first_band(const Lat &L, const Ext &E, const Part &Pt, bool valid, int x, int y,
                                           int z)
{
    __shared__ int bl[PB][6];
    __shared__ int nm;
    const int p = blockIdx.y;
    const int g[3] = {x, y, z}, n3[3] = {L.nx, L.ny, L.nz};
    const Box bp = box_of(Pt.center + 3 * p, Pt.rad[p], 3.0);
    bool first = valid;
    for (int q0 = 0; q0 < p; q0 += PB) {
        __syncthreads();
        if (threadIdx.x == 0) nm = 0;
        __syncthreads();
        const int q = q0 + (int)threadIdx.x;
        if (q < p) {
            const Box b = box_of(Pt.center + 3 * q, Pt.rad[q], 3.0);
            int lo[3];
            bool meet = true;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const bool per = !bounded(E, a);
                lo[a] = per ? wrapn(b.lo[a], n3[a]) : b.lo[a];
                // the two boxes meet along the axis (periodic: some image does)
                const int d = per ? wrapn(lo[a] - bp.lo[a], n3[a]) : lo[a] - bp.lo[a];
                meet = meet && (per ? (d < bp.w[a] || d + b.w[a] > n3[a]) : (d < bp.w[a] && d + b.w[a] > 0));
            }
            if (meet) {
                const int j = atomicAdd(&nm, 1);
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    bl[j][a] = lo[a];
                    bl[j][3 + a] = b.w[a];
                }
            }
        }
        __syncthreads();
        const int cnt = nm;
        for (int j = 0; j < cnt && first; ++j) {
            bool in = true;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                int d = g[a] - bl[j][a];
                if (!bounded(E, a) && d < 0) d += n3[a]; // g and the wrapped corner lie in [0, n)
                in = in && d >= 0 && d < bl[j][3 + a];
            }
            if (in) first = false;
        }
    }
    return first;
}

// Node (x, y, z) of the band (index n, mask nt) in the scatter step, as the
// tile kernel updates its lanes, with the phase field and the normals read
// from memory (ext_normal: zero at wall ghosts and solid nodes, wetted at
// link nodes, S:536-594), refilled nodes from their equilibria (S:1364-1391)
// and the band's gradients and normals stored for the particle kernels.  Its
// terms go to the fixed-point sums of their destinations by atomic addition
// (exact and order-free): into a wall ghost they return to the node, into a
// solid node they are dropped (the link's bounce-back value replaces them,
// k_curve_fs).  Solid nodes pass their pulled totals through (S:622, S:712,
// S:743) and add no terms.
template <class S>
__device__ __forceinline__ void sc_band_node(const Lat &L, const Ext &E, const S *__restrict__ fin,
                                             S *__restrict__ fout, float4 *__restrict__ sout,
                                             const float *__restrict__ phi, const float *__restrict__ nrm,
                                             int2 *__restrict__ own, int x, int y, int z, unsigned n, unsigned nt)
{
    using R = float;
    const size_t N = (size_t)L.N;
    const unsigned sx = (unsigned)L.ny * (unsigned)L.nz;
    const R rho0 = R(L.rho0), qs = SC_ONE / R(L.rho0);
    auto idx = [&](int dx, int dy, int dz) {
        return (unsigned)wrapn(x + dx, L.nx) * sx + (unsigned)wrapn(y + dy, L.ny) * (unsigned)L.nz +
               (unsigned)wrapn(z + dz, L.nz);
    };
    R gx, gy, gz;
    grad19<R>([&](int dx, int dy, int dz) { return phi_of_q(at(phi, idx(dx, dy, dz))); }, gx, gy, gz);
    const R g2 = gx * gx + gy * gy + gz * gz;
    bool flag = g2 > R(EVAP_LIM * EVAP_LIM);
    // the normals of the node and its neighbours (k_band_normals)
    auto n_at = [&](int k, int dx, int dy, int dz) { return at(nrm + (size_t)k * N, idx(dx, dy, dz)); };
    const R nx = n_at(0, 0, 0, 0), ny = n_at(1, 0, 0, 0), nz = n_at(2, 0, 0, 0);
    if (nt & NT_SOLID) flag = false;
    if (flag && (nt & NT_POUT)) { // wetting rescales the gradient too (S:569-581)
        const R fnorm = fsqrt(g2);
        gx = fnorm * nx;
        gy = fnorm * ny;
        gz = fnorm * nz;
    }
    const bool force_on = flag && !(nt & NT_POUT) && away_from_walls(E, L, x, y, z);
    {
        // the band's gradient (zero at solid nodes) and normal
        const bool fluid = !(nt & NT_SOLID);
        E.der[n] = fluid ? double(gx) : 0.0;
        E.der[N + n] = fluid ? double(gy) : 0.0;
        E.der[2 * N + n] = fluid ? double(gz) : 0.0;
        E.der[3 * N + n] = double(nx);
        E.der[4 * N + n] = double(ny);
        E.der[5 * N + n] = double(nz);
    }
    unsigned xo[3], yo[3], zo[3];
    node_offsets(L, x, y, z, xo, yo, zo);
    // the wall ghost slots the node reads at the next step (as the tile kernel
    // writes them): slot i of the ghost x - c_i takes the node's population opp i
    const unsigned wbb = wall_bb_mask(E, L, x, y, z);
    auto ghost_slot = [&](int i, R v) {
        if ((wbb >> i) & 1u)
            StoreOf<S>::st(fout + (size_t)i * N, xo[1 - CX[i]] + yo[1 - CY[i]] + zo[1 - CZ[i]], v);
    };
    if (nt & NT_SOLID) {
        // each slot takes the total it pulls; its colours (side array) split
        // with the source node's colour fraction
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            const unsigned m = xo[1 - CX[i]] + yo[1 - CY[i]] + zo[1 - CZ[i]];
            const R f = StoreOf<S>::ld(fin + (size_t)i * N, m);
            StoreOf<S>::st(fout + (size_t)i * N, n, f);
            if (i > 0) ghost_slot(OPC[i], f);
            const float qm = at(phi, m);
            const R fa = f + R(W[i]) * rho0, hm = __builtin_fabsf(qm) * fa;
            at((float *)E.rbo[0] + (size_t)i * N, n) =
                __builtin_signbitf(qm) ? enc_slot(fa - hm, hm) : enc_slot(hm, fa - hm);
        }
        at(sout, n) = make_float4(FS_SOLID, 0.f, 0.f, 0.f);
        return;
    }
    R f[Q];
    Moments<R, true> mo;
    R hm;
    bool redmaj;
    if (nt & NT_REFILL) {
        // SRT equilibria of the refill densities and the rigid-body velocity
        const double rr = E.mac[N + n], rb = E.mac[2 * N + n];
        const R rt = R(rr + rb);
        const R ux = R(E.mac[3 * N + n]), uy = R(E.mac[4 * N + n]), uz = R(E.mac[5 * N + n]);
        const R usqr = ux * ux + uy * uy + uz * uz;
        R s = 0;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            const R uc = ux * R(CX[i]) + uy * R(CY[i]) + uz * R(CZ[i]);
            const R poly = R(1) + R(3) * uc + R(4.5) * uc * uc - R(1.5) * usqr;
            f[i] = R(W[i]) * (rt * poly - rho0);
            s += f[i];
        }
        mo.drho = s;
        mo.rho = R(E.mac[n]);
        mo.ux = ux;
        mo.uy = uy;
        mo.uz = uz;
        mo.sr = R(rr);
        mo.sb = R(rb);
        redmaj = rb <= rr;
        hm = R((redmaj ? rb : rr) / (rr + rb));
    } else {
        R df = 0, jx = 0, jy = 0, jz = 0;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            f[i] = StoreOf<S>::ld(fin + (size_t)i * N, xo[1 - CX[i]] + yo[1 - CY[i]] + zo[1 - CZ[i]]);
            df += f[i];
            if (CX[i] == 1) jx += f[i];
            if (CX[i] == -1) jx -= f[i];
            if (CY[i] == 1) jy += f[i];
            if (CY[i] == -1) jy -= f[i];
            if (CZ[i] == 1) jz += f[i];
            if (CZ[i] == -1) jz -= f[i];
        }
        mo.rho = rho0 + df;
        mo.drho = df;
        const R irho = fdiv(R(1), mo.rho);
        mo.ux = jx * irho;
        mo.uy = jy * irho;
        mo.uz = jz * irho;
        const float qn = at(phi, n);
        redmaj = __builtin_signbitf(qn);
        hm = __builtin_fabsf(qn);
        const R smin = mo.rho * hm;
        mo.sr = redmaj ? mo.rho - smin : smin;
        mo.sb = redmaj ? smin : mo.rho - smin;
    }
    const R phin = phi_of_q(at(phi, n));
    mo.phi = phin;
    // normal derivatives for the curvature (S:635-660)
    R d[3][3];
    if (force_on) curv19<R>(n_at, d);
    collide<R, true>(L, f, mo, phin, gx, gy, gz, nx, ny, nz, force_on, d);
    const Recol<R> rc = recol<R, true>(L, mo, flag);
    R g = rc.gam;
    R Mf = hm;
    if (rc.evr) { g = 0; Mf = redmaj ? R(1) : R(0); }
    if (rc.evb) { g = 0; Mf = redmaj ? R(0) : R(1); }
    const R Gx = g * nx, Gy = g * ny, Gz = g * nz;
    const R Mg = redmaj ? -g : g;
    // the next state's totals and recolour scalars (k_curve_fs reads those
    // of the link nodes)
#pragma unroll
    for (int i = 0; i < Q; ++i) StoreOf<S>::st(fout + (size_t)i * N, n, f[i]);
    if (wbb) {
        // (offsets formed again, as in the tile kernel)
        int yy = y, zz = z;
        asm volatile("" : "+v"(yy), "+v"(zz));
        unsigned gxo[3], gyo[3], gzo[3];
        node_offsets(L, x, yy, zz, gxo, gyo, gzo);
#pragma unroll
        for (int i = 1; i < Q; ++i)
            if ((wbb >> i) & 1u)
                StoreOf<S>::st(fout + (size_t)i * N, gxo[1 - CX[i]] + gyo[1 - CY[i]] + gzo[1 - CZ[i]], f[OPC[i]]);
    }
    {
        Recol<R> ce = rc;
        ce.M = hm;
        ce.redmaj = redmaj;
        at(sout, n) = make_float4(encode_frac(ce), float(Gx), float(Gy), float(Gz));
    }
    // the terms, formed as the tile kernel forms them
    const R s1 = R(W[1]) * qs, s7 = R(W[7]) * qs;
    const R Mgx = Mg * nx, Mgy = Mg * ny, Mgz = Mg * nz;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        R v = f[i];
        if constexpr (sizeof(S) != 4) v = R(S(v)); // the value the next step reads
        const R fq = __builtin_fmaf(v, qs, R(W[i] * double(SC_ONE)));
        R cq = Mf * fq;
        if (i > 0 && i <= 6) cq = __builtin_fmaf(Mf, fq, dot_c<R>(i, Mgx * s1, Mgy * s1, Mgz * s1));
        if (i > 6) cq = __builtin_fmaf(Mf, fq, dot_c<R>(i, Mgx * s7, Mgy * s7, Mgz * s7));
        const int qf = round_half_up(fq), qm = round_half_up(cq);
        const int2 c = redmaj ? make_int2(qf - qm, qm) : make_int2(qm, qf - qm);
        unsigned dst = n;
        if (i > 0 && !is_wall_ghost(E, L, wrapn(x + CX[i], L.nx), wrapn(y + CY[i], L.ny), wrapn(z + CZ[i], L.nz))) {
            dst = idx(CX[i], CY[i], CZ[i]);
            if (at(E.ntype, dst) & NT_SOLID) continue;
        }
        atomic_add2(own + dst, c);
    }
}

// The band is a shell of its box: pack the lanes that hold a band node
// (mine), in box order, onto the workgroup's first lanes, so that whole
// waves work and the rest return.  Every lane of the workgroup must call it;
// lane threadIdx.x receives the packed value in out and returns true if it
// has one.
__device__ __forceinline__ bool pack_lanes(bool mine, int4 v, int4 &out)
{
    __shared__ int4 slot[PB];
    __shared__ int wcount[PB / 64];
    const unsigned lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    const unsigned long long bal = __ballot(mine);
    if (lane == 0) wcount[wv] = __popcll(bal);
    __syncthreads();
    int base = 0, total = 0;
#pragma unroll
    for (int i = 0; i < PB / 64; ++i) {
        base += i < (int)wv ? wcount[i] : 0;
        total += wcount[i];
    }
    if (mine) slot[base + __popcll(bal & ((1ull << lane) - 1ull))] = v;
    __syncthreads();
    if ((int)threadIdx.x >= total) return false;
    out = slot[threadIdx.x];
    return true;
}

// The band's step in the scatter form (after the tile kernel, which has
// written every node's own-box sums).
template <class S>
__global__ __launch_bounds__(PB) void k_collide_sc_band(Lat L, Ext E, Part Pt, const S *__restrict__ fin,
                                                        S *__restrict__ fout, float4 *__restrict__ sout,
                                                        const float *__restrict__ phi, const float *__restrict__ nrm,
                                                        int2 *__restrict__ own)
{
    const int p = blockIdx.y;
    const Box b = box_of(Pt.center + 3 * p, Pt.rad[p], 3.0);
    const unsigned k = blockIdx.x * PB + threadIdx.x;
    int X[3] = {0, 0, 0};
    unsigned n = 0;
    bool valid = k < box_size(b);
    if (valid) {
        box_node(L, b, k, X, n);
        valid = box_in_domain(E, L, X);
    }
    const int x = wrapn(X[0], L.nx), y = wrapn(X[1], L.ny), z = wrapn(X[2], L.nz);
    // band nodes only (not the box's corners, not the wall ghost layers)
    unsigned nt = 0;
    if (valid) {
        nt = at(E.ntype, n);
        valid = (nt & NT_NEAR) && !is_wall_ghost(E, L, x, y, z);
    }
    const bool mine = first_band(L, E, Pt, valid, x, y, z);
    // the sums are integers, so the lane that adds a term changes no result
    int4 s;
    if (!pack_lanes(mine, make_int4(x, y, z, (int)nt), s)) return;
    const unsigned m = ((unsigned)s.x * L.ny + s.y) * L.nz + s.z;
    sc_band_node<S>(L, E, fin, fout, sout, phi, nrm, own, s.x, s.y, s.z, m, (unsigned)s.w);
}

// The unit normals (ext_normal) of every node of each particle's box with
// margin R + 4, which holds every neighbour of its band, into nrm (3 x N,
// component-major), for the band kernel's curvature.  Workgroups (block,
// particle); a node in several boxes is written with equal values.
static __global__ __launch_bounds__(PB) void k_band_normals(Lat L, Ext E, Part Pt, const float *__restrict__ phi,
                                                            float *__restrict__ nrm)
{
    const int p = blockIdx.y;
    const Box b = box_of(Pt.center + 3 * p, Pt.rad[p], 4.0);
    const unsigned k = blockIdx.x * PB + threadIdx.x;
    if (k >= box_size(b)) return;
    int X[3];
    unsigned n;
    box_node(L, b, k, X, n);
    if (!box_in_domain(E, L, X)) return;
    float nv[3];
    ext_normal<float, QPhi>(L, E, (const QPhi *)phi, wrapn(X[0], L.nx), wrapn(X[1], L.ny), wrapn(X[2], L.nz), nv);
    const size_t N = (size_t)L.N;
    at(nrm, n) = nv[0];
    at(nrm + N, n) = nv[1];
    at(nrm + 2 * N, n) = nv[2];
}

// The band's fields at the start of a scatter-form step (band_macros): rho
// and the velocity from the gathered totals, the colour densities from the
// node's signed minority fraction; at solid nodes rho0, zero velocity and
// phi = 0 (calmacros, S:784-789).
template <class S>
__global__ __launch_bounds__(PB) void k_band_sc(Lat L, Ext E, Part Pt, const S *__restrict__ fs, float *__restrict__ phi)
{
    const int p = blockIdx.y;
    const Box b = box_of(Pt.center + 3 * p, Pt.rad[p], 3.0);
    const unsigned k = blockIdx.x * PB + threadIdx.x;
    int X[3] = {0, 0, 0};
    unsigned n0 = 0, nt0 = 0;
    bool mine = k < box_size(b);
    if (mine) {
        box_node(L, b, k, X, n0);
        mine = box_in_domain(E, L, X);
    }
    if (mine) {
        nt0 = at(E.ntype, n0);
        mine = (nt0 & NT_NEAR) && !is_wall_ghost(E, L, wrapn(X[0], L.nx), wrapn(X[1], L.ny), wrapn(X[2], L.nz));
    }
    int4 s;
    if (!pack_lanes(mine, make_int4(wrapn(X[0], L.nx), wrapn(X[1], L.ny), wrapn(X[2], L.nz), (int)nt0), s)) return;
    {
        using R = float;
        unsigned xo[3], yo[3], zo[3];
        node_offsets(L, s.x, s.y, s.z, xo, yo, zo);
        const unsigned n = xo[1] + yo[1] + zo[1];
        const unsigned nt = (unsigned)s.w;
        const size_t N = (size_t)L.N;
        const R rho0 = R(L.rho0);
        Moments<R, true> mo;
        if (nt & NT_SOLID) {
            at((QPhi *)phi, n) = QPhi(0.0f);
            mo.rho = rho0;
            mo.ux = mo.uy = mo.uz = 0;
            mo.sr = mo.sb = 0;
        } else {
            R df = 0, jx = 0, jy = 0, jz = 0;
#pragma unroll
            for (int i = 0; i < Q; ++i) {
                const R f = StoreOf<S>::ld(fs + (size_t)i * N, xo[1 - CX[i]] + yo[1 - CY[i]] + zo[1 - CZ[i]]);
                df += f;
                if (CX[i] == 1) jx += f;
                if (CX[i] == -1) jx -= f;
                if (CY[i] == 1) jy += f;
                if (CY[i] == -1) jy -= f;
                if (CZ[i] == 1) jz += f;
                if (CZ[i] == -1) jz -= f;
            }
            mo.rho = rho0 + df;
            const R irho = fdiv(R(1), mo.rho);
            mo.ux = jx * irho;
            mo.uy = jy * irho;
            mo.uz = jz * irho;
            const float qn = at(phi, n);
            const R smin = mo.rho * __builtin_fabsf(qn);
            mo.sr = __builtin_signbitf(qn) ? mo.rho - smin : smin;
            mo.sb = __builtin_signbitf(qn) ? smin : mo.rho - smin;
        }
        band_macros(L, E, n, nt, mo);
    }
}

} // namespace cgl
