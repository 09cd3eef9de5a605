/*
 * cgl_oracle.c -- FP64 CPU restatement of the reference's per-timestep
 * lattice update.  TEST INFRASTRUCTURE ONLY (see cgl_oracle.h): the product
 * never links or calls this file.
 *
 * Every routine cites the reference lines it restates as S:N, meaning
 * Source1.f90:N of X-Sheen/Symmetry.  The arithmetic keeps the reference's
 * operation order (Fortran left-to-right evaluation) so that, built with
 * -ffp-contract=off, results agree with the compiled reference to a few ulp.
 * Loops keep the reference's x -> y -> z nesting so that order-dependent
 * reductions (momentum, capillary force) sum in the same order as a
 * single-threaded reference run.
 */
#include "cgl_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define Q 19
#define EVAP_LIM 1.0e-8 /* S:16 */

/* D3Q19 velocity set in the reference's direction order (S:132-134). */
static const int CX[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
static const int CY[Q] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
static const int CZ[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};

static double WI[Q];
static int OPC[Q], SYM[3][Q];
static double MM[Q][Q], MINV[Q][Q];
static int tables_ready = 0;

static int find_dir(int cx, int cy, int cz)
{
    for (int j = 0; j < Q; ++j)
        if (CX[j] == cx && CY[j] == cy && CZ[j] == cz) return j;
    return -1;
}

/* Rows of the d'Humieres D3Q19 moment matrix as polynomials of c (the
 * reference stores the same matrix as a DATA table, S:105-123). */
static double moment_row(int k, int i)
{
    double cx = CX[i], cy = CY[i], cz = CZ[i];
    double c2 = cx * cx + cy * cy + cz * cz;
    switch (k) {
    case 0: return 1.0;
    case 1: return 19.0 * c2 - 30.0;
    case 2: return (21.0 * c2 * c2 - 53.0 * c2 + 24.0) / 2.0;
    case 3: return cx;
    case 4: return (5.0 * c2 - 9.0) * cx;
    case 5: return cy;
    case 6: return (5.0 * c2 - 9.0) * cy;
    case 7: return cz;
    case 8: return (5.0 * c2 - 9.0) * cz;
    case 9: return 3.0 * cx * cx - c2;
    case 10: return (3.0 * c2 - 5.0) * (3.0 * cx * cx - c2);
    case 11: return cy * cy - cz * cz;
    case 12: return (3.0 * c2 - 5.0) * (cy * cy - cz * cz);
    case 13: return cx * cy;
    case 14: return cy * cz;
    case 15: return cx * cz;
    case 16: return cx * (cy * cy - cz * cz);
    case 17: return cy * (cz * cz - cx * cx);
    default: return cz * (cx * cx - cy * cy);
    }
}

static void init_tables(void)
{
    if (tables_ready) return;
    for (int i = 0; i < Q; ++i) {
        int c2 = CX[i] * CX[i] + CY[i] * CY[i] + CZ[i] * CZ[i];
        WI[i] = c2 == 0 ? 1.0 / 3.0 : (c2 == 1 ? 1.0 / 18.0 : 1.0 / 36.0); /* S:140-142 */
        OPC[i] = find_dir(-CX[i], -CY[i], -CZ[i]);                           /* S:135 */
        SYM[0][i] = find_dir(-CX[i], CY[i], CZ[i]);                          /* S:136 */
        SYM[1][i] = find_dir(CX[i], -CY[i], CZ[i]);                          /* S:137 */
        SYM[2][i] = find_dir(CX[i], CY[i], -CZ[i]);                          /* S:138 */
    }
    for (int k = 0; k < Q; ++k)
        for (int i = 0; i < Q; ++i) MM[k][i] = moment_row(k, i);
    /* M_inv(i,j) = M(j,i) / <M(j,:),M(j,:)> (S:126-130). */
    for (int i = 0; i < Q; ++i)
        for (int j = 0; j < Q; ++j) {
            double nrm = 0.0;
            for (int k = 0; k < Q; ++k) nrm += MM[j][k] * MM[j][k];
            MINV[i][j] = MM[j][i] / nrm;
        }
    tables_ready = 1;
}

void or_tables(double *wi, int *c, int *opc, int *sym, double *M, double *Minv)
{
    init_tables();
    for (int i = 0; i < Q; ++i) {
        wi[i] = WI[i];
        c[i] = CX[i];
        c[Q + i] = CY[i];
        c[2 * Q + i] = CZ[i];
        opc[i] = OPC[i];
        for (int a = 0; a < 3; ++a) sym[a * Q + i] = SYM[a][i];
    }
    for (int k = 0; k < Q; ++k)
        for (int i = 0; i < Q; ++i) {
            M[k * Q + i] = MM[k][i];
            Minv[k * Q + i] = MINV[k][i];
        }
}

/* ---------------------------------------------------------------- helpers */

static inline size_t IDX(const or_state *s, int x, int y, int z)
{
    return ((size_t)x * s->ny + y) * s->nz + z;
}

static inline int wrap(int v, int n)
{
    return v < 0 ? v + n : (v >= n ? v - n : v);
}

/* next(i, x) of the periodic neighbour tables (S:295-324). */
static inline size_t NB(const or_state *s, int x, int y, int z, int i)
{
    return IDX(s, wrap(x + CX[i], s->nx), wrap(y + CY[i], s->ny), wrap(z + CZ[i], s->nz));
}

static inline size_t NBM(const or_state *s, int x, int y, int z, int i)
{
    return IDX(s, wrap(x - CX[i], s->nx), wrap(y - CY[i], s->ny), wrap(z - CZ[i], s->nz));
}

static void interior(const or_state *s, int lo[3], int hi[3])
{
    int n[3] = {s->nx, s->ny, s->nz};
    for (int a = 0; a < 3; ++a) {
        lo[a] = s->bc[2 * a] == OR_BC_PERIODIC ? 0 : 1;
        hi[a] = s->bc[2 * a + 1] == OR_BC_PERIODIC ? n[a] - 1 : n[a] - 2;
    }
}

#define FOR_INTERIOR(s, x, y, z)                                                   \
    int lo_[3], hi_[3];                                                            \
    interior(s, lo_, hi_);                                                         \
    for (int x = lo_[0]; x <= hi_[0]; ++x)                                         \
        for (int y = lo_[1]; y <= hi_[1]; ++y)                                     \
            for (int z = lo_[2]; z <= hi_[2]; ++z)

/* Plane copy on one face: dst plane (ghost) <- src plane (adjacent interior).
 * kind 0: scalar copy, 1: vector with the normal component negated,
 * 2: populations mirrored with sym_axis. */
static void face_planes(const or_state *s, int axis, int side, int *dst, int *src)
{
    int n = axis == 0 ? s->nx : (axis == 1 ? s->ny : s->nz);
    *dst = side ? n - 1 : 0;
    *src = side ? n - 2 : 1;
}

static void plane_loop_dims(const or_state *s, int axis, int *n1, int *n2)
{
    int n[3] = {s->nx, s->ny, s->nz};
    *n1 = n[(axis + 1) % 3];
    *n2 = n[(axis + 2) % 3];
}

static inline size_t PIDX(const or_state *s, int axis, int p, int u, int v)
{
    int c[3];
    c[axis] = p;
    c[(axis + 1) % 3] = u;
    c[(axis + 2) % 3] = v;
    return IDX(s, c[0], c[1], c[2]);
}

static void copy_scalar_face(const or_state *s, double *f, int axis, int side)
{
    int d, sr, n1, n2;
    face_planes(s, axis, side, &d, &sr);
    plane_loop_dims(s, axis, &n1, &n2);
    for (int u = 0; u < n1; ++u)
        for (int v = 0; v < n2; ++v) f[PIDX(s, axis, d, u, v)] = f[PIDX(s, axis, sr, u, v)];
}

static void copy_int_face(const or_state *s, int *f, int axis, int side)
{
    int d, sr, n1, n2;
    face_planes(s, axis, side, &d, &sr);
    plane_loop_dims(s, axis, &n1, &n2);
    for (int u = 0; u < n1; ++u)
        for (int v = 0; v < n2; ++v) f[PIDX(s, axis, d, u, v)] = f[PIDX(s, axis, sr, u, v)];
}

static void copy_vec_face(const or_state *s, double *vx, double *vy, double *vz, int axis, int side)
{
    int d, sr, n1, n2;
    double *v3[3] = {vx, vy, vz};
    face_planes(s, axis, side, &d, &sr);
    plane_loop_dims(s, axis, &n1, &n2);
    for (int u = 0; u < n1; ++u)
        for (int v = 0; v < n2; ++v) {
            size_t a = PIDX(s, axis, d, u, v), b = PIDX(s, axis, sr, u, v);
            for (int k = 0; k < 3; ++k) v3[k][a] = (k == axis) ? -v3[k][b] : v3[k][b];
        }
}

static void copy_pop_face(const or_state *s, double *f, int axis, int side)
{
    int d, sr, n1, n2;
    face_planes(s, axis, side, &d, &sr);
    plane_loop_dims(s, axis, &n1, &n2);
    for (int u = 0; u < n1; ++u)
        for (int v = 0; v < n2; ++v) {
            size_t a = PIDX(s, axis, d, u, v) * Q, b = PIDX(s, axis, sr, u, v) * Q;
            for (int i = 0; i < Q; ++i) f[a + i] = f[b + SYM[axis][i]];
        }
}

/* Mirror ghost fills, axis order x, y, z and high face before low face, as
 * symmetry_{X,Y,Z}_{phi,u,fi} do (S:1443-1542). */
static void mirror_scalar(const or_state *s, double *f)
{
    for (int a = 0; a < 3; ++a)
        for (int side = 1; side >= 0; --side)
            if (s->bc[2 * a + side] == OR_BC_MIRROR) copy_scalar_face(s, f, a, side);
}

static void mirror_vec(const or_state *s, double *vx, double *vy, double *vz)
{
    for (int a = 0; a < 3; ++a)
        for (int side = 1; side >= 0; --side)
            if (s->bc[2 * a + side] == OR_BC_MIRROR) copy_vec_face(s, vx, vy, vz, a, side);
}

static void mirror_pops(const or_state *s, double *f)
{
    for (int a = 0; a < 3; ++a)
        for (int side = 1; side >= 0; --side)
            if (s->bc[2 * a + side] == OR_BC_MIRROR) copy_pop_face(s, f, a, side);
}

/* calmacros / refilling mirror block: every field for x, then y, then z
 * (S:809-828, S:1397-1416). */
static void mirror_macros(or_state *s, double *popf)
{
    for (int a = 0; a < 3; ++a)
        for (int side = 1; side >= 0; --side) {
            if (s->bc[2 * a + side] != OR_BC_MIRROR) continue;
            copy_scalar_face(s, s->phi, a, side);
            copy_scalar_face(s, s->rho_r, a, side);
            copy_scalar_face(s, s->rho_b, a, side);
            copy_scalar_face(s, s->rho, a, side);
            copy_vec_face(s, s->ux, s->uy, s->uz, a, side);
            copy_pop_face(s, popf, a, side);
        }
}

/* ------------------------------------------------------------ equilibria */

/* equilf_SRT (S:833-848). */
void or_equilf_srt(double *feq, double rho, double ux, double uy, double uz)
{
    init_tables();
    double usqr = ux * ux + uy * uy + uz * uz;
    for (int i = 0; i < Q; ++i) {
        double uc = ux * CX[i] + uy * CY[i] + uz * CZ[i];
        feq[i] = rho * WI[i] * (1.0 + 3.0 * uc + 4.5 * uc * uc - 1.5 * usqr);
    }
}

/* equilf: equilibrium moments and Guo forcing moments (S:397-450). */
void or_equilf(double *meq, double *fk, double rho, double ux, double uy, double uz,
               double fx, double fy, double fz)
{
    double usqr = ux * ux + uy * uy + uz * uz;
    meq[0] = 1.0;
    meq[1] = 19.0 * usqr - 11.0;
    meq[2] = 3.0 - 5.5 * usqr;
    meq[3] = ux;
    meq[4] = -ux / 1.5;
    meq[5] = uy;
    meq[6] = -uy / 1.5;
    meq[7] = uz;
    meq[8] = -uz / 1.5;
    meq[9] = 3.0 * ux * ux - usqr;
    meq[10] = -0.5 * (3.0 * ux * ux - usqr);
    meq[11] = uy * uy - uz * uz;
    meq[12] = -0.5 * (uy * uy - uz * uz);
    meq[13] = ux * uy;
    meq[14] = uy * uz;
    meq[15] = ux * uz;
    meq[16] = 0.0;
    meq[17] = 0.0;
    meq[18] = 0.0;
    for (int k = 0; k < Q; ++k) meq[k] = rho * meq[k];
    double uf = ux * fx + uy * fy + uz * fz;
    fk[0] = 0.0;
    fk[1] = 38.0 * uf;
    fk[2] = -11.0 * uf;
    fk[3] = fx;
    fk[4] = -fx / 1.5;
    fk[5] = fy;
    fk[6] = -fy / 1.5;
    fk[7] = fz;
    fk[8] = -fz / 1.5;
    fk[9] = 6.0 * ux * fx - 2.0 * uf;
    fk[10] = -(3.0 * ux * fx - uf);
    fk[11] = 2.0 * (uy * fy - uz * fz);
    fk[12] = -(uy * fy - uz * fz);
    fk[13] = ux * fy + uy * fx;
    fk[14] = uy * fz + uz * fy;
    fk[15] = ux * fz + uz * fx;
    fk[16] = 0.0;
    fk[17] = 0.0;
    fk[18] = 0.0;
}

/* ------------------------------------------------------------ fluid stages */

/* calmacros (S:770-831): moments at interior nodes, wall-layer phi copies
 * (S:806-807, z then x then y), then the mirror fills. */
void or_calmacros(or_state *s)
{
    init_tables();
    FOR_INTERIOR(s, x, y, z)
    {
        size_t n = IDX(s, x, y, z);
        if (s->walls_all[n] >= 1) {
            s->phi[n] = 0.0;
            s->ux[n] = 0.0;
            s->uy[n] = 0.0;
            s->uz[n] = 0.0;
            s->rho[n] = s->rho0;
        } else {
            double *f = s->fi + n * Q, *r = s->ri + n * Q, *b = s->bi + n * Q;
            double sr = 0.0, sb = 0.0, jx = 0.0, jy = 0.0, jz = 0.0;
            for (int i = 0; i < Q; ++i) f[i] = r[i] + b[i];
            for (int i = 0; i < Q; ++i) sr += r[i];
            for (int i = 0; i < Q; ++i) sb += b[i];
            s->rho_r[n] = sr;
            s->rho_b[n] = sb;
            s->rho[n] = sr + sb;
            for (int i = 0; i < Q; ++i) jx += f[i] * CX[i];
            for (int i = 0; i < Q; ++i) jy += f[i] * CY[i];
            for (int i = 0; i < Q; ++i) jz += f[i] * CZ[i];
            s->ux[n] = jx / s->rho[n];
            s->uy[n] = jy / s->rho[n];
            s->uz[n] = jz / s->rho[n];
            s->phi[n] = (sr - sb) / s->rho[n];
        }
    }
    static const int order[3] = {2, 0, 1};
    for (int k = 0; k < 3; ++k) {
        int a = order[k];
        for (int side = 0; side < 2; ++side)
            if (s->bc[2 * a + side] == OR_BC_WALL) copy_scalar_face(s, s->phi, a, side);
    }
    mirror_macros(s, s->fi);
}

/* derivatives (S:452-597): ghost phi at live solid nodes, mirror fills,
 * phase-field gradient and unit normal, Akai wetting correction at live fluid
 * nodes, mirror fills of n and grad. */
void or_derivatives(or_state *s)
{
    init_tables();
    /* ghost phi, S:468-492 */
    {
        FOR_INTERIOR(s, x, y, z)
        {
            size_t n = IDX(s, x, y, z);
            if (s->pin[n] > 0) {
                double wsum = 0.0, acc = 0.0;
                for (int i = 1; i < Q; ++i) {
                    size_t m = NB(s, x, y, z, i);
                    if (s->walls_all[m] <= 0) {
                        wsum = wsum + WI[i];
                        acc = acc + WI[i] * s->phi[m];
                    }
                }
                s->phi[n] = acc / wsum;
            }
        }
    }
    mirror_scalar(s, s->phi); /* S:493-495 */
    /* gradient and normal, S:498-533 */
    {
        FOR_INTERIOR(s, x, y, z)
        {
            size_t n = IDX(s, x, y, z);
            s->gx[n] = 0.0;
            s->gy[n] = 0.0;
            s->gz[n] = 0.0;
            s->nxv[n] = 0.0;
            s->nyv[n] = 0.0;
            s->nzv[n] = 0.0;
            s->flag[n] = 0;
            if (s->walls_all[n] <= 0) {
                double gx = 0.0, gy = 0.0, gz = 0.0;
                for (int i = 1; i < Q; ++i) {
                    double val = WI[i] * s->phi[NB(s, x, y, z, i)];
                    gx = gx + val * CX[i];
                    gy = gy + val * CY[i];
                    gz = gz + val * CZ[i];
                }
                gx = 3.0 * gx;
                gy = 3.0 * gy;
                gz = 3.0 * gz;
                s->gx[n] = gx;
                s->gy[n] = gy;
                s->gz[n] = gz;
                double fnorm = sqrt(gx * gx + gy * gy + gz * gz);
                if (fnorm > EVAP_LIM) {
                    s->nxv[n] = gx / fnorm;
                    s->nyv[n] = gy / fnorm;
                    s->nzv[n] = gz / fnorm;
                    s->flag[n] = 1;
                }
            }
        }
    }
    /* wetting, S:536-587 */
    {
        FOR_INTERIOR(s, x, y, z)
        {
            size_t n = IDX(s, x, y, z);
            if (s->pout[n] >= 1 && s->flag[n]) {
                int pp = s->pout[n] - 1;
                double nw[3] = {(double)(x + 1) - s->center[3 * pp], (double)(y + 1) - s->center[3 * pp + 1],
                                (double)(z + 1) - s->center[3 * pp + 2]};
                double st = sqrt(nw[0] * nw[0] + nw[1] * nw[1] + nw[2] * nw[2]);
                nw[0] /= st;
                nw[1] /= st;
                nw[2] /= st;
                double nt[3] = {s->nxv[n], s->nyv[n], s->nzv[n]};
                double dmax = fabs(nw[0] - nt[0]);
                if (fabs(nw[1] - nt[1]) > dmax) dmax = fabs(nw[1] - nt[1]);
                if (fabs(nw[2] - nt[2]) > dmax) dmax = fabs(nw[2] - nt[2]);
                double theta = dmax < 1.0e-5 ? 1.0e-5 : acos(nw[0] * nt[0] + nw[1] * nt[1] + nw[2] * nt[2]);
                double ca = cos(s->cita[pp]), sa = sin(s->cita[pp]);
                double a1 = ca - sa * cos(theta) / sin(theta);
                double a2 = ca + sa * cos(theta) / sin(theta);
                double b1 = sa / sin(theta);
                double b2 = -sa / sin(theta);
                double v1[3], v2[3];
                for (int k = 0; k < 3; ++k) {
                    v1[k] = a1 * nw[k] + b1 * nt[k];
                    v2[k] = a2 * nw[k] + b2 * nt[k];
                }
                double db1 = (v1[0] - nt[0]) * (v1[0] - nt[0]) + (v1[1] - nt[1]) * (v1[1] - nt[1]) +
                             (v1[2] - nt[2]) * (v1[2] - nt[2]);
                double db2 = (v2[0] - nt[0]) * (v2[0] - nt[0]) + (v2[1] - nt[1]) * (v2[1] - nt[1]) +
                             (v2[2] - nt[2]) * (v2[2] - nt[2]);
                double *v = db1 < db2 ? v1 : v2;
                s->nxv[n] = v[0];
                s->nyv[n] = v[1];
                s->nzv[n] = v[2];
                double fnorm = sqrt(s->gx[n] * s->gx[n] + s->gy[n] * s->gy[n] + s->gz[n] * s->gz[n]);
                s->gx[n] = fnorm * v[0];
                s->gy[n] = fnorm * v[1];
                s->gz[n] = fnorm * v[2];
            }
        }
    }
    mirror_vec(s, s->nxv, s->nyv, s->nzv); /* S:589-591 */
    mirror_vec(s, s->gx, s->gy, s->gz);    /* S:592-594 */
}

/* Force guard of S:624 generalised: no curvature force next to a wall ghost
 * layer, whose normals the gradient pass never writes. */
static int away_from_walls(const or_state *s, int x, int y, int z)
{
    int c[3] = {x, y, z}, n[3] = {s->nx, s->ny, s->nz};
    for (int a = 0; a < 3; ++a) {
        if (s->bc[2 * a] == OR_BC_WALL && c[a] < 2) return 0;
        if (s->bc[2 * a + 1] == OR_BC_WALL && c[a] > n[a] - 3) return 0;
    }
    return 1;
}

/* collision (S:599-695): curvature, CSF force, half-force velocity shift,
 * harmonic viscosity, MRT relaxation with Guo forcing; copy to ffi. */
void or_collision(or_state *s)
{
    init_tables();
    FOR_INTERIOR(s, x, y, z)
    {
        size_t n = IDX(s, x, y, z);
        if (s->walls_all[n] > 0) continue;
        double force[3] = {0.0, 0.0, 0.0};
        if (s->flag[n] && s->pout[n] <= 0 && away_from_walls(s, x, y, z)) {
            double d[3][3] = {{0}};
            for (int i = 1; i < Q; ++i) {
                size_t m = NB(s, x, y, z, i);
                double nv[3] = {s->nxv[m], s->nyv[m], s->nzv[m]};
                for (int k = 0; k < 3; ++k) {
                    double tmp = WI[i] * nv[k];
                    d[k][0] = d[k][0] + tmp * CX[i];
                    d[k][1] = d[k][1] + tmp * CY[i];
                    d[k][2] = d[k][2] + tmp * CZ[i];
                }
            }
            for (int k = 0; k < 3; ++k)
                for (int l = 0; l < 3; ++l) d[k][l] = 3.0 * d[k][l];
            double nx = s->nxv[n], ny = s->nyv[n], nz = s->nzv[n];
            /* d[k][l] = d n_k / d x_l */
            double kurv = nx * ny * (d[1][0] + d[0][1]) + nx * nz * (d[2][0] + d[0][2]) +
                          ny * nz * (d[1][2] + d[2][1]) - (ny * ny + nz * nz) * d[0][0] -
                          (nx * nx + nz * nz) * d[1][1] - (nx * nx + ny * ny) * d[2][2];
            double val = 0.5 * s->sigma * kurv;
            force[0] = val * s->gx[n];
            force[1] = val * s->gy[n];
            force[2] = val * s->gz[n];
        }
        s->ux[n] = s->ux[n] + 0.5 * force[0] / s->rho[n];
        s->uy[n] = s->uy[n] + 0.5 * force[1] / s->rho[n];
        s->uz[n] = s->uz[n] + 0.5 * force[2] / s->rho[n];
        double ph = s->phi[n], nu;
        if (ph <= -1.0)
            nu = s->nu_b;
        else if (ph >= 1.0)
            nu = s->nu_r;
        else
            nu = 2.0 / ((1.0 + ph) / s->nu_r + (1.0 - ph) / s->nu_b);
        double su = 1.0 / (3.0 * nu + 0.5);
        double sq = 8.0 * (2.0 - su) / (8.0 - su);
        double si[Q] = {1.0, su, su, 1.0, sq, 1.0, sq, 1.0, sq, su, su, su, su, su, su, su, sq, sq, sq};
        double *f = s->fi + n * Q;
        double mi[Q], meq[Q], fk[Q];
        for (int k = 0; k < Q; ++k) {
            double acc = 0.0;
            for (int j = 0; j < Q; ++j) acc += MM[k][j] * f[j];
            mi[k] = acc;
        }
        or_equilf(meq, fk, s->rho[n], s->ux[n], s->uy[n], s->uz[n], force[0], force[1], force[2]);
        for (int k = 0; k < Q; ++k) mi[k] = mi[k] - si[k] * (mi[k] - meq[k]) + (1.0 - 0.5 * si[k]) * fk[k];
        for (int i = 0; i < Q; ++i) {
            double acc = 0.0;
            for (int k = 0; k < Q; ++k) acc += MINV[i][k] * mi[k];
            f[i] = acc;
            s->ffi[n * Q + i] = acc;
        }
    }
}

/* Recolor (S:697-732): Latva-Kokko recolouring with the EVAP_LIM cases. */
void or_recolor(or_state *s)
{
    init_tables();
    FOR_INTERIOR(s, x, y, z)
    {
        size_t n = IDX(s, x, y, z);
        if (s->walls_all[n] >= 1) continue;
        double *f = s->fi + n * Q, *r = s->ri + n * Q, *b = s->bi + n * Q;
        double rr = s->rho_r[n], rb = s->rho_b[n], rho = s->rho[n];
        for (int i = 0; i < Q; ++i) r[i] = rr / rho * f[i];
        if (s->flag[n]) {
            double nx = s->nxv[n], ny = s->nyv[n], nz = s->nzv[n];
            for (int i = 1; i < Q; ++i)
                r[i] = r[i] + rr * rb / rho * s->beta * WI[i] * (nx * CX[i] + ny * CY[i] + nz * CZ[i]);
        }
        if (rr <= EVAP_LIM)
            for (int i = 0; i < Q; ++i) r[i] = 0.0;
        if (rb <= EVAP_LIM)
            for (int i = 0; i < Q; ++i) r[i] = f[i];
        for (int i = 0; i < Q; ++i) b[i] = f[i] - r[i];
    }
}

/* One wall face of boundary(): half-way bounce-back of the populations that
 * leave the interior-adjacent plane into the ghost plane (S:1557-1590). */
static void bounce_face(or_state *s, double *r, double *b, int axis, int side)
{
    int src, ghost, n1, n2;
    face_planes(s, axis, side, &ghost, &src);
    plane_loop_dims(s, axis, &n1, &n2);
    const int *cc = axis == 0 ? CX : (axis == 1 ? CY : CZ);
    int want = side ? 1 : -1;
    for (int u = 0; u < n1; ++u)
        for (int v = 0; v < n2; ++v) {
            int c[3];
            c[axis] = src;
            c[(axis + 1) % 3] = u;
            c[(axis + 2) % 3] = v;
            size_t n = IDX(s, c[0], c[1], c[2]);
            for (int i = 1; i < Q; ++i) {
                if (cc[i] != want) continue;
                size_t m = NB(s, c[0], c[1], c[2], i);
                r[m * Q + OPC[i]] = r[n * Q + i];
                b[m * Q + OPC[i]] = b[n * Q + i];
            }
        }
}

/* boundary (S:1544-1592): mirror fills of r and b, then wall bounce-back
 * (z faces, then x faces, then y faces). */
void or_boundary(or_state *s)
{
    init_tables();
    for (int a = 0; a < 3; ++a)
        for (int side = 1; side >= 0; --side)
            if (s->bc[2 * a + side] == OR_BC_MIRROR) {
                copy_pop_face(s, s->ri, a, side);
                copy_pop_face(s, s->bi, a, side);
            }
    static const int order[3] = {2, 0, 1};
    for (int k = 0; k < 3; ++k) {
        int a = order[k];
        for (int side = 0; side < 2; ++side)
            if (s->bc[2 * a + side] == OR_BC_WALL) bounce_face(s, s->ri, s->bi, a, side);
    }
}

/* streaming (S:735-768): periodic pull of directions 1..18 over all nodes. */
void or_streaming(or_state *s, double *f)
{
    init_tables();
    size_t N = (size_t)s->nx * s->ny * s->nz;
    double *ft = (double *)malloc(N * Q * sizeof(double));
    for (int x = 0; x < s->nx; ++x)
        for (int y = 0; y < s->ny; ++y)
            for (int z = 0; z < s->nz; ++z) {
                size_t n = IDX(s, x, y, z);
                ft[n * Q] = f[n * Q];
                for (int i = 1; i < Q; ++i) ft[n * Q + i] = f[NBM(s, x, y, z, i) * Q + i];
            }
    memcpy(f, ft, N * Q * sizeof(double));
    free(ft);
}

/* --------------------------------------------------------- particle stages */

static inline void wall_velocity(const or_state *s, int pp, const double *ctr, int x, int y, int z,
                                 double *xr, double uw[3])
{
    xr[0] = (double)(x + 1) - ctr[3 * pp];
    xr[1] = (double)(y + 1) - ctr[3 * pp + 1];
    xr[2] = (double)(z + 1) - ctr[3 * pp + 2];
    const double *u = s->up + 3 * pp, *w = s->rp + 3 * pp;
    uw[0] = u[0] + xr[2] * w[1] - xr[1] * w[2];
    uw[1] = u[1] + xr[0] * w[2] - xr[2] * w[0];
    uw[2] = u[2] + xr[1] * w[0] - xr[0] * w[1];
}

/* curve (S:990-1034): moving-wall half-way bounce-back on links from live
 * fluid nodes into their particle, then mirror fills of r and b. */
void or_curve(or_state *s)
{
    init_tables();
    if (s->npar > 0) {
        FOR_INTERIOR(s, x, y, z)
        {
            size_t n = IDX(s, x, y, z);
            if (s->pout[n] < 1) continue;
            int pp = s->pout[n] - 1;
            for (int i = 1; i < Q; ++i) {
                size_t m = NBM(s, x, y, z, i);
                if (s->walls[m] != pp + 1) continue;
                double xr[3], uw[3];
                wall_velocity(s, pp, s->center, x, y, z, xr, uw);
                double star = uw[0] * CX[i] + uw[1] * CY[i] + uw[2] * CZ[i];
                s->ri[n * Q + i] = s->ri[m * Q + OPC[i]] + 6.0 * WI[i] * s->rho_r[n] * star;
                s->bi[n * Q + i] = s->bi[m * Q + OPC[i]] + 6.0 * WI[i] * s->rho_b[n] * star;
            }
        }
    }
    mirror_pops(s, s->ri);
    mirror_pops(s, s->bi);
}

/* set_walls (S:326-395): solid masks by the FP64 distance test, live fluid /
 * live solid marks, mirror-face copies.  Serial particle order: the highest
 * particle index wins where spheres or their neighbourhoods overlap. */
void or_set_walls(or_state *s)
{
    init_tables();
    size_t N = (size_t)s->nx * s->ny * s->nz;
    memcpy(s->walls_old, s->walls_all, N * sizeof(int));
    memset(s->walls_all, 0, N * sizeof(int));
    memset(s->walls, 0, N * sizeof(int));
    memset(s->pin, 0, N * sizeof(int));
    memset(s->pout, 0, N * sizeof(int));
    for (int pp = 0; pp < s->npar; ++pp) {
        double cx = s->center[3 * pp], cy = s->center[3 * pp + 1], cz = s->center[3 * pp + 2];
        double r2 = s->rad[pp] * s->rad[pp];
        for (int x = 0; x < s->nx; ++x)
            for (int y = 0; y < s->ny; ++y)
                for (int z = 0; z < s->nz; ++z) {
                    double dx = (double)(x + 1) - cx, dy = (double)(y + 1) - cy, dz = (double)(z + 1) - cz;
                    double dist = dx * dx + dy * dy + dz * dz;
                    if (dist <= r2) {
                        size_t n = IDX(s, x, y, z);
                        s->walls[n] = pp + 1;
                        s->walls_all[n] = 10;
                    }
                }
    }
    for (int pp = 0; pp < s->npar; ++pp) {
        FOR_INTERIOR(s, x, y, z)
        {
            size_t n = IDX(s, x, y, z);
            if (s->walls_all[n] != 0) continue;
            for (int i = 1; i < Q; ++i) {
                size_t m = NB(s, x, y, z, i);
                if (s->walls[m] == pp + 1) {
                    s->pout[n] = pp + 1;
                    s->pin[m] = pp + 1;
                }
            }
        }
    }
    for (int a = 0; a < 3; ++a)
        for (int side = 1; side >= 0; --side)
            if (s->bc[2 * a + side] == OR_BC_MIRROR) {
                copy_int_face(s, s->walls, a, side);
                copy_int_face(s, s->walls_all, a, side);
                copy_int_face(s, s->pin, a, side);
                copy_int_face(s, s->pout, a, side);
            }
}

/* refilling (S:1351-1418): nodes uncovered by a moving particle get the rigid
 * velocity, neighbour-mean colour densities and SRT equilibria. */
void or_refilling(or_state *s)
{
    init_tables();
    FOR_INTERIOR(s, x, y, z)
    {
        size_t n = IDX(s, x, y, z);
        if (!(s->pout[n] >= 1 && s->walls_old[n] >= 1)) continue;
        int pp = s->pout[n] - 1;
        double xr[3], uw[3];
        wall_velocity(s, pp, s->center_old, x, y, z, xr, uw);
        s->ux[n] = uw[0];
        s->uy[n] = uw[1];
        s->uz[n] = uw[2];
        double sr = 0.0, sb = 0.0, num = 0.0;
        for (int i = 1; i < Q; ++i) {
            size_t m = NB(s, x, y, z, i);
            if (s->walls_old[m] <= 0) {
                sr = sr + s->rho_r[m];
                sb = sb + s->rho_b[m];
                num = num + 1.0;
            }
        }
        s->rho_r[n] = sr / num;
        s->rho_b[n] = sb / num;
        s->rho[n] = s->rho_r[n] + s->rho_b[n];
        s->phi[n] = (s->rho_r[n] - s->rho_b[n]) / s->rho[n];
        or_equilf_srt(s->ri + n * Q, s->rho_r[n], uw[0], uw[1], uw[2]);
        or_equilf_srt(s->bi + n * Q, s->rho_b[n], uw[0], uw[1], uw[2]);
        for (int i = 0; i < Q; ++i) s->fi[n * Q + i] = s->ri[n * Q + i] + s->bi[n * Q + i];
    }
    mirror_macros(s, s->fi);
}

/* momentum (S:1178-1261): Galilean-invariant momentum exchange on the links
 * of each particle; force and torque on the particle. */
void or_momentum(or_state *s)
{
    init_tables();
    for (int k = 0; k < 3 * s->npar; ++k) {
        s->fp[k] = 0.0;
        s->tp[k] = 0.0;
    }
    FOR_INTERIOR(s, x, y, z)
    {
        size_t n = IDX(s, x, y, z);
        if (s->pout[n] < 1) continue;
        int pp = s->pout[n] - 1;
        for (int i = 1; i < Q; ++i) {
            size_t m = NBM(s, x, y, z, i);
            if (s->walls[m] != pp + 1) continue;
            double xr[3], uw[3];
            wall_velocity(s, pp, s->center, x, y, z, xr, uw);
            int o = OPC[i];
            double fi = s->fi[n * Q + i], fo = s->ffi[n * Q + o];
            double fmx = ((double)CX[i] - uw[0]) * fi - ((double)CX[o] - uw[0]) * fo;
            double fmy = ((double)CY[i] - uw[1]) * fi - ((double)CY[o] - uw[1]) * fo;
            double fmz = ((double)CZ[i] - uw[2]) * fi - ((double)CZ[o] - uw[2]) * fo;
            double tx = xr[1] * fmz - xr[2] * fmy;
            double ty = xr[2] * fmx - xr[0] * fmz;
            double tz = xr[0] * fmy - xr[1] * fmx;
            s->fp[3 * pp] = s->fp[3 * pp] - fmx;
            s->fp[3 * pp + 1] = s->fp[3 * pp + 1] - fmy;
            s->fp[3 * pp + 2] = s->fp[3 * pp + 2] - fmz;
            s->tp[3 * pp] = s->tp[3 * pp] - tx;
            s->tp[3 * pp + 1] = s->tp[3 * pp + 1] - ty;
            s->tp[3 * pp + 2] = s->tp[3 * pp + 2] - tz;
        }
    }
}

static void unitize(double a[3])
{
    double st = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
    st = sqrt(st);
    a[0] = a[0] / st;
    a[1] = a[1] / st;
    a[2] = a[2] / st;
}

static void vcross(const double a[3], const double b[3], double c[3])
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = -(a[0] * b[2] - a[2] * b[0]);
    c[2] = a[0] * b[1] - a[1] * b[0];
}

/* Lagrange_points + capillary_force2 (S:1799-2028): 180 x 91 surface points
 * per sphere, IDW-interpolated phase field, wetting-corrected normal and the
 * capillary force and torque on each sphere.  The per-step Laglange.dat
 * write (S:1968-1974) is output, not state, and is not restated. */
void or_lagrange_capillary(or_state *s)
{
    init_tables();
    static const int ex[8] = {0, 1, 1, 0, 0, 1, 1, 0};
    static const int ey[8] = {0, 0, 1, 1, 0, 0, 1, 1};
    static const int ez[8] = {0, 0, 0, 0, 1, 1, 1, 1};
    const double PI = 4.0 * atan(1.0);
    const double alpha = 2.0 / 180.0 * PI;
    const int La_m = (int)(2.0 * PI / alpha), La_n = (int)(PI / alpha);
    const double kcap = 0.134 * 1.43;
    int nn[3] = {s->nx, s->ny, s->nz};
    for (int pp = 0; pp < s->npar; ++pp) {
        s->shb[pp] = 0.0;
        for (int k = 0; k < 3; ++k) {
            s->tension[3 * pp + k] = 0.0;
            s->tension_t[3 * pp + k] = 0.0;
        }
    }
    for (int pp = 0; pp < s->npar; ++pp) {
        const double *c = s->center + 3 * pp;
        double R = s->rad[pp];
        double cita = s->cita[pp];
        for (int na = 1; na <= La_m; ++na)
            for (int nb = 1; nb <= La_n + 1; ++nb) {
                double aa = (double)(na - 1) * alpha, bb = (double)(nb - 1) * alpha;
                double L[3];
                L[0] = c[0] + R * sin(bb) * cos(aa);
                L[1] = c[1] + R * sin(bb) * sin(aa);
                L[2] = c[2] + R * cos(bb);
                double ds = R * R * sin(bb) * (alpha * alpha);
                double lrho = 0.0;
                int lflag = 0, in = 1, c0[3] = {0, 0, 0};
                for (int a = 0; a < 3; ++a) {
                    if (s->dom_lo_on[a] && (s->dom_lo_strict[a] ? !(L[a] > s->dom_lo[a]) : !(L[a] >= s->dom_lo[a])))
                        in = 0;
                    if (s->dom_hi_on[a] && !(L[a] < s->dom_hi[a])) in = 0;
                }
                if (in) {
                    for (int a = 0; a < 3; ++a) c0[a] = (int)L[a];
                    double wsum = 0.0;
                    for (int k = 0; k < 8; ++k) {
                        int f1[3] = {c0[0] + ex[k], c0[1] + ey[k], c0[2] + ez[k]};
                        size_t m = IDX(s, wrap(f1[0] - 1, nn[0]), wrap(f1[1] - 1, nn[1]), wrap(f1[2] - 1, nn[2]));
                        if (s->walls[m] <= 0) {
                            double star = s->phi[m] + s->gx[m] * (L[0] - (double)f1[0]) +
                                          s->gy[m] * (L[1] - (double)f1[1]) + s->gz[m] * (L[2] - (double)f1[2]);
                            double dd = ((double)f1[0] - L[0]) * ((double)f1[0] - L[0]) +
                                        ((double)f1[1] - L[1]) * ((double)f1[1] - L[1]) +
                                        ((double)f1[2] - L[2]) * ((double)f1[2] - L[2]);
                            double wgt = 1.0 / dd;
                            lrho = lrho + wgt * star;
                            wsum = wsum + wgt;
                        }
                    }
                    lrho = lrho / wsum;
                    if (lrho < -1.0) lrho = -1.0;
                    if (lrho > 1.0) lrho = 1.0;
                    if (fabs(lrho) < 0.99) lflag = 1;
                }
                if (lrho > 0.0) s->shb[pp] = s->shb[pp] + ds;
                if (!lflag) continue;
                double nw[3] = {L[0] - c[0], L[1] - c[1], L[2] - c[2]};
                unitize(nw);
                double dmin = 100.0;
                int best[3] = {c0[0], c0[1], c0[2]};
                for (int k = 0; k < 8; ++k) {
                    int f1[3] = {c0[0] + ex[k], c0[1] + ey[k], c0[2] + ez[k]};
                    size_t m = IDX(s, wrap(f1[0] - 1, nn[0]), wrap(f1[1] - 1, nn[1]), wrap(f1[2] - 1, nn[2]));
                    if (s->walls[m] > 0) continue;
                    if (!(fabs(s->nxv[m] * s->nxv[m] + s->nyv[m] * s->nyv[m] + s->nzv[m] * s->nzv[m]) > 0.0))
                        continue;
                    double dd = ((double)f1[0] - L[0]) * ((double)f1[0] - L[0]) +
                                ((double)f1[1] - L[1]) * ((double)f1[1] - L[1]) +
                                ((double)f1[2] - L[2]) * ((double)f1[2] - L[2]);
                    if (sqrt(dd) < dmin) {
                        dmin = sqrt(dd);
                        best[0] = f1[0];
                        best[1] = f1[1];
                        best[2] = f1[2];
                    }
                }
                size_t mb = IDX(s, wrap(best[0] - 1, nn[0]), wrap(best[1] - 1, nn[1]), wrap(best[2] - 1, nn[2]));
                double nt[3] = {s->nxv[mb], s->nyv[mb], s->nzv[mb]};
                unitize(nt);
                double theta = acos(nw[0] * nt[0] + nw[1] * nt[1] + nw[2] * nt[2]);
                double ca = cos(cita), sa = sin(cita);
                double a1 = ca - sa * cos(theta) / sin(theta);
                double a2 = ca + sa * cos(theta) / sin(theta);
                double b1 = sa / sin(theta), b2 = -sa / sin(theta);
                double v1[3], v2[3];
                for (int k = 0; k < 3; ++k) {
                    v1[k] = a1 * nw[k] + b1 * nt[k];
                    v2[k] = a2 * nw[k] + b2 * nt[k];
                }
                double db1 = (v1[0] - nt[0]) * (v1[0] - nt[0]) + (v1[1] - nt[1]) * (v1[1] - nt[1]) +
                             (v1[2] - nt[2]) * (v1[2] - nt[2]);
                double db2 = (v2[0] - nt[0]) * (v2[0] - nt[0]) + (v2[1] - nt[1]) * (v2[1] - nt[1]) +
                             (v2[2] - nt[2]) * (v2[2] - nt[2]);
                double ln[3];
                for (int k = 0; k < 3; ++k) ln[k] = db1 < db2 ? v1[k] : v2[k];
                /* capillary_force2, S:1995-2022 */
                double xr[3] = {L[0] - c[0], L[1] - c[1], L[2] - c[2]};
                double nw2[3] = {xr[0], xr[1], xr[2]};
                unitize(nw2);
                double ie[3] = {-ln[0], -ln[1], -ln[2]};
                double tmp[3], ntan[3];
                vcross(ie, nw2, tmp);
                double proj = sqrt(tmp[0] * tmp[0] + tmp[1] * tmp[1] + tmp[2] * tmp[2]);
                vcross(tmp, ie, ntan);
                unitize(ntan);
                double om = 1.0 - lrho * lrho;
                double star = 4.5 * kcap * s->beta * (om * om) * ds * proj;
                double fv[3];
                int right = cita == PI / 2.0;
                for (int k = 0; k < 3; ++k) fv[k] = star * (right ? nw2[k] : ntan[k]) * s->sigma;
                double tt[3];
                vcross(xr, fv, tt);
                for (int k = 0; k < 3; ++k) {
                    s->tension[3 * pp + k] = s->tension[3 * pp + k] + fv[k];
                    s->tension_t[3 * pp + k] = s->tension_t[3 * pp + k] + tt[k];
                }
            }
    }
}

/* repulsive (S:1594-1671): soft wall and pair repulsion, range 3. */
void or_repulsive(or_state *s)
{
    const double range = 3.0, order = 0.01;
    for (int k = 0; k < 3 * s->npar; ++k) s->frep[k] = 0.0;
    for (int pp = 0; pp < s->npar; ++pp) {
        double R = s->rad[pp];
        /* S:1633-1650: high faces push with +, low faces with - */
        for (int a = 0; a < 3; ++a) {
            if (s->rep_on[2 * a]) {
                double dlo = s->center[3 * pp + a] - s->rep_plane[2 * a] - R;
                if (dlo < range) s->frep[3 * pp + a] -= 2.0 * (s->fg / order) * ((dlo - range) * (dlo - range)) / (range * range);
            }
            if (s->rep_on[2 * a + 1]) {
                double dhi = s->rep_plane[2 * a + 1] - s->center[3 * pp + a] - R;
                if (dhi < range) s->frep[3 * pp + a] += 2.0 * (s->fg / order) * ((dhi - range) * (dhi - range)) / (range * range);
            }
        }
    }
    for (int i = 0; i < s->npar; ++i)
        for (int j = 0; j < s->npar; ++j) {
            if (i == j) continue;
            const double *ci = s->center + 3 * i, *cj = s->center + 3 * j;
            double d = sqrt((ci[0] - cj[0]) * (ci[0] - cj[0]) + (ci[1] - cj[1]) * (ci[1] - cj[1]) +
                            (ci[2] - cj[2]) * (ci[2] - cj[2])) -
                       (s->rad[i] + s->rad[j]);
            if (d < range)
                for (int k = 0; k < 3; ++k)
                    s->frep[3 * i + k] =
                        s->frep[3 * i + k] - (s->fg / order) * ((d - range) * (d - range)) / (range * range) * (ci[k] - cj[k]) / d;
        }
}

/* movement (S:1263-1349): reference 1/8-symmetry rules (only with
 * ref_rules), then the explicit Newton-Euler update. */
void or_movement(or_state *s)
{
    int np = s->npar;
    if (s->ref_rules) {
        double xl = s->nx, yl = s->ny, zl = s->nz;
        for (int p = 0; p < np; ++p) {
            double *F = s->fp + 3 * p, *T = s->tp + 3 * p, *Fr = s->frep + 3 * p;
            double *te = s->tension + 3 * p, *tt = s->tension_t + 3 * p;
            const double *c0 = s->center0 + 3 * p;
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
        }
        int p = np / 3; /* pp == par/3+1 in 1-based numbering, S:1322 */
        if (p < np) {
            double *F = s->fp + 3 * p, *T = s->tp + 3 * p, *Fr = s->frep + 3 * p;
            double *te = s->tension + 3 * p, *tt = s->tension_t + 3 * p;
            double t = (te[0] + te[2]) / 2.0;
            te[0] = t; te[2] = t;
            tt[1] = 0.0;
            t = (F[0] + F[2]) / 2.0;
            F[0] = t; F[2] = t;
            T[1] = 0.0;
            t = (Fr[0] + Fr[2]) / 2.0;
            Fr[0] = t; Fr[2] = t;
        }
    }
    for (int p = 0; p < np; ++p)
        for (int k = 0; k < 3; ++k) {
            double ftot = s->fp[3 * p + k] + s->frep[3 * p + k] + s->tension[3 * p + k];
            s->up[3 * p + k] = s->up[3 * p + k] + ftot / s->mass[p];
            s->rp[3 * p + k] = s->rp[3 * p + k] + (s->tp[3 * p + k] + s->tension_t[3 * p + k]) / s->inertia[p];
        }
    memcpy(s->center_old, s->center, 3 * np * sizeof(double));
    for (int p = 0; p < np; ++p)
        for (int k = 0; k < 3; ++k) {
            double ftot = s->fp[3 * p + k] + s->frep[3 * p + k] + s->tension[3 * p + k];
            s->center[3 * p + k] = s->center[3 * p + k] + s->up[3 * p + k] + 0.5 * ftot / s->mass[p];
        }
}

/* One time step in the reference's stage order (S:216-251), output block
 * excluded. */
void or_step(or_state *s, int particles)
{
    or_calmacros(s);
    if (particles && s->npar > 0) {
        or_lagrange_capillary(s);
        or_momentum(s);
        or_repulsive(s);
        or_movement(s);
        or_set_walls(s);
        or_refilling(s);
    }
    or_derivatives(s);
    or_collision(s);
    or_recolor(s);
    or_boundary(s);
    or_streaming(s, s->ri);
    or_streaming(s, s->bi);
    or_curve(s);
}

void or_run(or_state *s, int nsteps, int particles)
{
    for (int t = 0; t < nsteps; ++t) or_step(s, particles);
}
