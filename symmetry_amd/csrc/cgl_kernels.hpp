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
//   K1 k_moments  pull-gather -> rho_R, rho_B -> phi          (calmacros, S:770-831)
//   K2 k_collide  x-marching tile, LDS rings of phi and n    (derivatives S:452-597,
//                 -> grad phi, n, curvature, CSF force,        collision S:599-695,
//                 MRT + Guo, recolour, write pops              Recolor S:697-732,
//                                                              streaming S:735-768)
//
// FP64 ("plain") keeps the oracle's exact operation order and is compiled
// without FMA contraction, so it reproduces the FP64 CPU path bit for bit.
// FP32/FP16 storage keeps each population shifted by w_i*rho0*a of the node
// that will read it, a = (1 + phi)/2 for red and (1 - phi)/2 for blue, with
// phi of the step that wrote it: the writer has that phi in its LDS ring and
// the reader needs only its own node's value.  Collision works on the
// deviations f - w_i*rho0 in FP32 with compensated moment sums.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cgl_lattice.hpp"

namespace cgl {

constexpr double EVAP_LIM = 1.0e-8; // S:16
constexpr int TZ = 64;              // workgroup extent along z (one wavefront)
constexpr int TY = 4;               // workgroup extent along y

struct Lat {
    int nx, ny, nz;
    int xcd_remap; // 1: hand each XCD a contiguous range of tiles
    long long N;
    double rho0, sigma, beta, nu_r, nu_b;
};

template <class S> struct StoreOf;
template <> struct StoreOf<double> {
    __device__ static double ld(const double *p) { return *p; }
    __device__ static void st(double *p, double v) { *p = v; }
};
template <> struct StoreOf<float> {
    __device__ static float ld(const float *p) { return *p; }
    __device__ static void st(float *p, float v) { *p = v; }
};
template <> struct StoreOf<_Float16> {
    __device__ static float ld(const _Float16 *p) { return (float)*p; }
    __device__ static void st(_Float16 *p, float v) { *p = (_Float16)v; }
};

__device__ __forceinline__ int wrap1(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }
__device__ __forceinline__ int wrapm(int v, int n)
{
    v %= n;
    return v < 0 ? v + n : v;
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
    const unsigned full = nb - nb % NXCD;
    if (L.xcd_remap && lin < full) lin = (lin % NXCD) * (full / NXCD) + lin / NXCD;
    bx = lin % gx;
    by = (lin / gx) % gy;
    bz = lin / (gx * gy);
}

// Neumaier compensated sum.
template <class R> struct CSum {
    R s = R(0), c = R(0);
    __device__ __forceinline__ void add(R x)
    {
        R t = s + x;
        if (__builtin_fabs(s) >= __builtin_fabs(x))
            c += (s - t) + x;
        else
            c += (x - t) + s;
        s = t;
    }
    __device__ __forceinline__ R get() const { return s + c; }
};

// Pull gather of the 19 red and blue populations at node (x, y, z), with
// unshifting.  off[d][k] is the wrapped index offset of coordinate d at
// displacement k-1.
template <class S, class R, class P, bool SHIFT>
__device__ __forceinline__ void gather(const Lat &L, const S *__restrict__ rs, const S *__restrict__ bs,
                                       const int xo[3], const int yo[3], const int zo[3], R r[Q], R b[Q])
{
    const size_t N = (size_t)L.N;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        int m = xo[1 - CX[i]] + yo[1 - CY[i]] + zo[1 - CZ[i]];
        r[i] = StoreOf<S>::ld(rs + (size_t)i * N + m);
        b[i] = StoreOf<S>::ld(bs + (size_t)i * N + m);
    }
}

// Moments of the gathered populations.
//  FP64 path: calmacros (S:792-799) operation for operation.
//  Shifted path: r, b hold deviations from w_i*rho0*ax and w_i*rho0*(1-ax),
//  ax = (1 + phi_old(x))/2 of this (reading) node.
template <class R, bool SHIFT> struct Moments {
    R sr, sb, rho, drho, ux, uy, uz, phi;
};

template <class R, bool SHIFT>
__device__ __forceinline__ Moments<R, SHIFT> moments(const Lat &L, const R r[Q], const R b[Q], R ax)
{
    Moments<R, SHIFT> m;
    if constexpr (!SHIFT) {
        R f[Q];
#pragma unroll
        for (int i = 0; i < Q; ++i) f[i] = r[i] + b[i];
        R sr = 0, sb = 0;
#pragma unroll
        for (int i = 0; i < Q; ++i) sr += r[i];
#pragma unroll
        for (int i = 0; i < Q; ++i) sb += b[i];
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
        m.drho = m.rho - R(L.rho0);
        m.ux = jx / m.rho;
        m.uy = jy / m.rho;
        m.uz = jz / m.rho;
        m.phi = (sr - sb) / m.rho;
    } else {
        CSum<R> cr, cb, cf;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            cr.add(r[i]);
            cb.add(b[i]);
            cf.add(r[i] + b[i]);
        }
        const R rho0 = R(L.rho0);
        m.sr = cr.get() + rho0 * ax;
        m.sb = cb.get() + rho0 * (R(1) - ax);
        m.drho = cf.get();
        m.rho = rho0 + m.drho;
        R jx = 0, jy = 0, jz = 0;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            R f = r[i] + b[i];
            if (CX[i] == 1) jx += f;
            if (CX[i] == -1) jx -= f;
            if (CY[i] == 1) jy += f;
            if (CY[i] == -1) jy -= f;
            if (CZ[i] == 1) jz += f;
            if (CZ[i] == -1) jz -= f;
        }
        m.ux = jx / m.rho;
        m.uy = jy / m.rho;
        m.uz = jz / m.rho;
        m.phi = (m.sr - m.sb) / (m.sr + m.sb);
    }
    return m;
}

// Node offsets (already multiplied by strides) for coordinates c-1, c, c+1.
__device__ __forceinline__ void node_offsets(const Lat &L, int x, int y, int z, int xo[3], int yo[3], int zo[3])
{
    const int sx = L.ny * L.nz;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        xo[k] = wrap1(x + k - 1, L.nx) * sx;
        yo[k] = wrap1(y + k - 1, L.ny) * L.nz;
        zo[k] = wrap1(z + k - 1, L.nz);
    }
}

// ---------------------------------------------------------------- K1 moments
// Writes phi of the step.  With FULL it also writes the macroscopic fields
// rho, rho_R, rho_B, ux, uy, uz, phi (bare velocity, as calmacros) as FP64.
template <class S, class R, class P, bool SHIFT, bool FULL>
__global__ __launch_bounds__(256) void k_moments(Lat L, const S *__restrict__ rs, const S *__restrict__ bs,
                                                 const P *__restrict__ phio, P *__restrict__ phi,
                                                 double *__restrict__ mac)
{
    unsigned bx, by, bz;
    tile_of_block(L, bx, by, bz);
    const int z = bx * TZ + threadIdx.x;
    const int y = by * TY + threadIdx.y;
    const int x = bz;
    if (z >= L.nz || y >= L.ny) return;
    int xo[3], yo[3], zo[3];
    node_offsets(L, x, y, z, xo, yo, zo);
    R r[Q], b[Q];
    gather<S, R, P, SHIFT>(L, rs, bs, xo, yo, zo, r, b);
    const int n = xo[1] + yo[1] + zo[1];
    R ax = 0;
    if constexpr (SHIFT) ax = R(0.5) * (R(1) + R(phio[n]));
    Moments<R, SHIFT> mo = moments<R, SHIFT>(L, r, b, ax);
    phi[n] = P(mo.phi);
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

// ----------------------------------------------------- gradient of phi (b)
// Gradient 3*sum_i w_i c_i phi(x + c_i) over i = 1..18 in the reference's
// order (S:511-520), with zero terms skipped (they add exact zeros).
template <class R, class F>
__device__ __forceinline__ void grad19(F phi_at, R &gx, R &gy, R &gz)
{
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

// ------------------------------------------------------------- collision (a)
// CSF force from the curvature of the normals (S:624-667), velocity shift,
// viscosity interpolation and MRT relaxation with Guo forcing (S:669-689),
// recolouring (S:708-726).  In: gathered pops, moments, phi, grad phi, n,
// flag, the 3x3 normal derivatives, and (shifted path) aw[i], the red shift
// fraction of the node that reads direction i.  Out: post-recolour pops
// (FP64) or their deviations from w_i*rho0*aw[i] / w_i*rho0*(1-aw[i]).
template <class R, bool SHIFT>
__device__ __forceinline__ void collide_recolor(const Lat &L, const R r[Q], const R b[Q], const Moments<R, SHIFT> &mo,
                                                R phin, R gx, R gy, R gz, R nx, R ny, R nz, bool flag,
                                                bool force_on, const R d[3][3], const R aw[Q], R rout[Q],
                                                R bout[Q])
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
    const R ux = mo.ux + R(0.5) * fx / rho;
    const R uy = mo.uy + R(0.5) * fy / rho;
    const R uz = mo.uz + R(0.5) * fz / rho;
    R nu;
    if (phin <= R(-1))
        nu = R(L.nu_b);
    else if (phin >= R(1))
        nu = R(L.nu_r);
    else
        nu = R(2) / ((R(1) + phin) / R(L.nu_r) + (R(1) - phin) / R(L.nu_b));
    const R su = R(1) / (R(3) * nu + R(0.5));
    const R sq = R(8) * (R(2) - su) / (R(8) - su);
    R f[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) f[i] = r[i] + b[i];
    R m[Q];
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
    // equilibrium and forcing moments (equilf, S:405-447)
    R meq[Q], fk[Q];
    {
        R usqr = ux * ux + uy * uy + uz * uz;
        meq[0] = 1;
        meq[1] = R(19) * usqr - R(11);
        meq[2] = R(3) - R(5.5) * usqr;
        meq[3] = ux;
        meq[4] = -ux / R(1.5);
        meq[5] = uy;
        meq[6] = -uy / R(1.5);
        meq[7] = uz;
        meq[8] = -uz / R(1.5);
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
        fk[4] = -fx / R(1.5);
        fk[5] = fy;
        fk[6] = -fy / R(1.5);
        fk[7] = fz;
        fk[8] = -fz / R(1.5);
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
    // recolouring (S:708-726)
    const R sr = mo.sr, sb = mo.sb;
    if constexpr (!SHIFT) {
#pragma unroll
        for (int i = 0; i < Q; ++i) rout[i] = sr / rho * f[i];
        if (flag) {
#pragma unroll
            for (int i = 1; i < Q; ++i) {
                R dot = 0;
                bool first = true;
                if (CX[i] != 0) { dot = CX[i] == 1 ? nx : -nx; first = false; }
                if (CY[i] != 0) { R t = CY[i] == 1 ? ny : -ny; dot = first ? t : dot + t; first = false; }
                if (CZ[i] != 0) { R t = CZ[i] == 1 ? nz : -nz; dot = first ? t : dot + t; }
                rout[i] = rout[i] + sr * sb / rho * R(L.beta) * R(W[i]) * dot;
            }
        }
        if (sr <= R(EVAP_LIM)) {
#pragma unroll
            for (int i = 0; i < Q; ++i) rout[i] = 0;
        }
        if (sb <= R(EVAP_LIM)) {
#pragma unroll
            for (int i = 0; i < Q; ++i) rout[i] = f[i];
        }
#pragma unroll
        for (int i = 0; i < Q; ++i) bout[i] = f[i] - rout[i];
    } else {
        // f holds deviations f' - w rho0; write r' - w rho0 aw, b' - w rho0 (1 - aw)
        const R rho0 = R(L.rho0);
        const R ar = sr / (sr + sb);
        const R gam = flag ? sr * sb / (sr + sb) * R(L.beta) : R(0);
        const bool evr = sr <= R(EVAP_LIM), evb = sb <= R(EVAP_LIM);
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            R dot = R(CX[i]) * nx + R(CY[i]) * ny + R(CZ[i]) * nz;
            const R a = aw[i];
            R v = ar * f[i] + R(W[i]) * rho0 * (ar - a) + gam * R(W[i]) * dot;
            if (evr) v = -R(W[i]) * rho0 * a;
            if (evb) v = f[i] + R(W[i]) * rho0 * (R(1) - a);
            rout[i] = v;
            bout[i] = f[i] - v;
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
template <class R, class P> struct K2Smem {
    P phi[5][TY + 4][TZ + 4];
    R n[3][3][TY + 2][TZ + 2];
};

__device__ __forceinline__ int slot5(int p) { return (p + 10) % 5; }
__device__ __forceinline__ int slot3(int p) { return (p + 9) % 3; }

template <class S, class R, class P, bool SHIFT>
__global__ __launch_bounds__(256) void k_collide(Lat L, const S *__restrict__ rin, const S *__restrict__ bin,
                                                 S *__restrict__ rout, S *__restrict__ bout,
                                                 const P *__restrict__ phi, const P *__restrict__ phio, int xc)
{
    __shared__ K2Smem<R, P> sm;
    const int tz = threadIdx.x, ty = threadIdx.y, tid = ty * TZ + tz;
    unsigned bx, by, bz;
    tile_of_block(L, bx, by, bz);
    const int z0 = bx * TZ, y0 = by * TY, x0 = bz * xc;
    const int x1 = min(x0 + xc, L.nx);
    const int z = z0 + tz, y = y0 + ty;
    const bool active = z < L.nz && y < L.ny;
    const int sx = L.ny * L.nz;

    auto load_phi = [&](int p) {
        const int gx = wrapm(p, L.nx) * sx;
        P(*dst)[TZ + 4] = sm.phi[slot5(p)];
        for (int e = tid; e < (TY + 4) * (TZ + 4); e += TZ * TY) {
            int jy = e / (TZ + 4), jz = e - jy * (TZ + 4);
            int gy = wrapm(y0 - 2 + jy, L.ny), gz = wrapm(z0 - 2 + jz, L.nz);
            dst[jy][jz] = phi[gx + gy * L.nz + gz];
        }
    };
    // unit normal of plane p on the (TY+2) x (TZ+2) halo region (S:498-532)
    auto normals = [&](int p) {
        const int s = slot3(p);
        for (int e = tid; e < (TY + 2) * (TZ + 2); e += TZ * TY) {
            int jy = e / (TZ + 2), jz = e - jy * (TZ + 2);
            R gx, gy, gz;
            grad19<R>([&](int dx, int dy, int dz) { return sm.phi[slot5(p + dx)][jy + 1 + dy][jz + 1 + dz]; }, gx,
                      gy, gz);
            R fnorm = __builtin_sqrt(gx * gx + gy * gy + gz * gz);
            R nx = 0, ny = 0, nz = 0;
            if (fnorm > R(EVAP_LIM)) {
                nx = gx / fnorm;
                ny = gy / fnorm;
                nz = gz / fnorm;
            }
            sm.n[0][s][jy][jz] = nx;
            sm.n[1][s][jy][jz] = ny;
            sm.n[2][s][jy][jz] = nz;
        }
    };

    for (int p = x0 - 2; p <= x0 + 1; ++p) load_phi(p);
    __syncthreads();
    normals(x0 - 1);
    normals(x0);

    for (int x = x0; x < x1; ++x) {
        __syncthreads(); // readers of the slots refilled below are done
        load_phi(x + 2);
        __syncthreads();
        normals(x + 1);
        __syncthreads();
        if (!active) continue;

        // gradient, normal and flag at the node (S:498-532)
        R gx, gy, gz;
        grad19<R>([&](int dx, int dy, int dz) { return sm.phi[slot5(x + dx)][ty + 2 + dy][tz + 2 + dz]; }, gx, gy,
                  gz);
        const R fnorm = __builtin_sqrt(gx * gx + gy * gy + gz * gz);
        const bool flag = fnorm > R(EVAP_LIM);
        const R nx = sm.n[0][slot3(x)][ty + 1][tz + 1];
        const R ny = sm.n[1][slot3(x)][ty + 1][tz + 1];
        const R nz = sm.n[2][slot3(x)][ty + 1][tz + 1];
        // normal derivatives for the curvature (S:635-660)
        R d[3][3];
        if (flag) {
#pragma unroll
            for (int k = 0; k < 3; ++k) d[k][0] = d[k][1] = d[k][2] = 0;
#pragma unroll
            for (int i = 1; i < Q; ++i) {
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    R tmp = R(W[i]) * sm.n[k][slot3(x + CX[i])][ty + 1 + CY[i]][tz + 1 + CZ[i]];
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

        // pull gather = streaming (S:743-753)
        int xo[3], yo[3], zo[3];
        node_offsets(L, x, y, z, xo, yo, zo);
        R r[Q], b[Q];
        gather<S, R, P, SHIFT>(L, rin, bin, xo, yo, zo, r, b);
        const int n = x * sx + y * L.nz + z;
        R ax = 0, aw[Q];
        if constexpr (SHIFT) {
            ax = R(0.5) * (R(1) + R(phio[n]));
            // shift of the slot read by x + c_i at the next step: that node's phi now
#pragma unroll
            for (int i = 0; i < Q; ++i)
                aw[i] = R(0.5) * (R(1) + R(sm.phi[slot5(x + CX[i])][ty + 2 + CY[i]][tz + 2 + CZ[i]]));
        }
        const Moments<R, SHIFT> mo = moments<R, SHIFT>(L, r, b, ax);
        const R phin = R(sm.phi[slot5(x)][ty + 2][tz + 2]);
        R ro[Q], bo[Q];
        collide_recolor<R, SHIFT>(L, r, b, mo, phin, gx, gy, gz, nx, ny, nz, flag, flag, d, aw, ro, bo);
        const size_t N = (size_t)L.N;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            StoreOf<S>::st(rout + (size_t)i * N + n, ro[i]);
            StoreOf<S>::st(bout + (size_t)i * N + n, bo[i]);
        }
    }
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
    int xo[3], yo[3], zo[3];
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

// Shift reference a of a stored slot, computed exactly as the readers do.
template <class P> __device__ __forceinline__ float shift_a(P ph, int red)
{
    float a = 0.5f * (1.0f + float(ph));
    return red ? a : 1.0f - a;
}

// Pack one colour: S_i(y) = P_i(y + c_i) - w_i rho0 a(y + c_i) (inverse of
// the periodic streaming permutation, so that the next pull gather returns P;
// the shift is that of the reading node y + c_i).
template <class S, class P, bool SHIFT>
__global__ __launch_bounds__(256) void k_pack(Lat L, const double *__restrict__ aos, S *__restrict__ soa,
                                              const P *__restrict__ phio, int red)
{
    const int z = blockIdx.x * TZ + threadIdx.x;
    const int y = blockIdx.y * TY + threadIdx.y;
    const int x = blockIdx.z;
    if (z >= L.nz || y >= L.ny) return;
    int xo[3], yo[3], zo[3];
    node_offsets(L, x, y, z, xo, yo, zo);
    const int n = xo[1] + yo[1] + zo[1];
    const size_t N = (size_t)L.N;
    for (int i = 0; i < Q; ++i) {
        int src = xo[1 + CX[i]] + yo[1 + CY[i]] + zo[1 + CZ[i]];
        double v = aos[(size_t)src * Q + i];
        if constexpr (SHIFT) v = v - W[i] * (L.rho0 * (double)shift_a(phio[src], red));
        StoreOf<S>::st(soa + (size_t)i * N + n, (decltype(StoreOf<S>::ld(soa)))v);
    }
}

// Unpack one colour: P_i(x) = S_i(x - c_i) + w_i rho0 a(x), as FP64 AoS.
template <class S, class P, bool SHIFT>
__global__ __launch_bounds__(256) void k_unpack(Lat L, const S *__restrict__ soa, double *__restrict__ aos,
                                                const P *__restrict__ phio, int red)
{
    const int z = blockIdx.x * TZ + threadIdx.x;
    const int y = blockIdx.y * TY + threadIdx.y;
    const int x = blockIdx.z;
    if (z >= L.nz || y >= L.ny) return;
    int xo[3], yo[3], zo[3];
    node_offsets(L, x, y, z, xo, yo, zo);
    const int n = xo[1] + yo[1] + zo[1];
    const size_t N = (size_t)L.N;
    for (int i = 0; i < Q; ++i) {
        int src = xo[1 - CX[i]] + yo[1 - CY[i]] + zo[1 - CZ[i]];
        double v = double(StoreOf<S>::ld(soa + (size_t)i * N + src));
        if constexpr (SHIFT) v = v + W[i] * (L.rho0 * (double)shift_a(phio[n], red));
        aos[(size_t)n * Q + i] = v;
    }
}

// ------------------------------------------------- device initial condition
// Liquid cylinder along z through (xc, yc) (1-based lattice coordinates),
// radius R(z) = R0 + sum_m a_m cos(2 pi m (z - 1) / nz), tanh profile
// phi = tanh(kprof (R(z) - r)) and Laplace density jump of the reference's
// initialization (S:871-889), fluid at rest, SRT equilibria.
struct CylIC {
    double xc, yc, r0, kprof;
    int nmodes;
    int pad;
    double amp[16];
};

__device__ __forceinline__ double cyl_phi(const Lat &L, const CylIC &ic, int x, int y, int z, double *Rz_out)
{
    const double PI = 3.14159265358979323846;
    double Rz = ic.r0;
    for (int m = 0; m < ic.nmodes; ++m) Rz += ic.amp[m] * cos(2.0 * PI * (m + 1) * (double)z / (double)L.nz);
    double dx = (double)(x + 1) - ic.xc, dy = (double)(y + 1) - ic.yc;
    *Rz_out = Rz;
    return tanh((Rz - sqrt(dx * dx + dy * dy)) * ic.kprof);
}

// Pass 1: phi of the initial state, the shift reference of each reader.
template <class P>
__global__ __launch_bounds__(256) void k_init_phi(Lat L, CylIC ic, P *__restrict__ phio)
{
    const int z = blockIdx.x * TZ + threadIdx.x;
    const int y = blockIdx.y * TY + threadIdx.y;
    const int x = blockIdx.z;
    if (z >= L.nz || y >= L.ny) return;
    double Rz;
    phio[((size_t)x * L.ny + y) * L.nz + z] = P(cyl_phi(L, ic, x, y, z, &Rz));
}

// Pass 2: node x's equilibria P_i(x) = w_i rho_c go to slot S_i(x - c_i),
// shifted by node x's own fraction (x reads that slot).
template <class S, class P, bool SHIFT>
__global__ __launch_bounds__(256) void k_init_pops(Lat L, CylIC ic, S *__restrict__ rs, S *__restrict__ bs,
                                                   const P *__restrict__ phio)
{
    const int z = blockIdx.x * TZ + threadIdx.x;
    const int y = blockIdx.y * TY + threadIdx.y;
    const int x = blockIdx.z;
    if (z >= L.nz || y >= L.ny) return;
    int xo[3], yo[3], zo[3];
    node_offsets(L, x, y, z, xo, yo, zo);
    const size_t N = (size_t)L.N;
    double Rz;
    double ph = cyl_phi(L, ic, x, y, z, &Rz);
    double rho = ph > 0.0 ? L.rho0 + L.sigma / Rz * 3.0 : L.rho0;
    double rr = rho * (ph + 1.0) / 2.0, rb = rho - rr;
    const P phs = P(ph); // the value pass 1 stored for this node
    for (int i = 0; i < Q; ++i) {
        int dst = xo[1 - CX[i]] + yo[1 - CY[i]] + zo[1 - CZ[i]];
        double vr = rr * W[i], vb = rb * W[i];
        if constexpr (SHIFT) {
            vr -= W[i] * (L.rho0 * (double)shift_a(phs, 1));
            vb -= W[i] * (L.rho0 * (double)shift_a(phs, 0));
        }
        StoreOf<S>::st(rs + (size_t)i * N + dst, (decltype(StoreOf<S>::ld(rs)))vr);
        StoreOf<S>::st(bs + (size_t)i * N + dst, (decltype(StoreOf<S>::ld(bs)))vb);
    }
}

} // namespace cgl
