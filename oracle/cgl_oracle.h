/*
 * cgl_oracle.h -- FP64 CPU restatement of the reference's per-timestep
 * colour-gradient D3Q19 MRT lattice update (Source1.f90, X-Sheen/Symmetry).
 *
 * TEST INFRASTRUCTURE ONLY.  This oracle is the checker the GPU library is
 * compared against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product (symmetry_amd/libcgl.so) never
 * links, loads or calls it.
 *
 * Layout follows the reference's Fortran column-major arrays with 0-based
 * indices: scalars are [x][y][z] (z fastest, S:148-166), populations are
 * [x][y][z][i] (direction i fastest, S:167-170).  Coordinates handed to the
 * particle routines are the reference's 1-based lattice coordinates (node
 * index + 1), as in S:344 and S:1009.
 *
 * Boundary faces are generalised: each of the six faces is periodic, a
 * half-way bounce-back wall with a ghost layer (the reference's x=1 and z=1,
 * S:806-807, S:1557-1590) or a specular mirror ghost layer (the reference's
 * x=xl, y=1, y=yl, z=zl, S:1443-1542).
 *
 * Pins: every stage is checked bit for bit against the compiled reference on
 * its own geometry (tests/test_oracle_vs_ref.py); periodic faces and
 * high-side walls, which the reference lacks, by translation and reflection
 * symmetry (tests/test_oracle_invariance.py).  Contact angles other than the
 * reference's compile-time 90 degrees (the wetting correction, S:536-587, and
 * the S:2015-2019 else-branch of the capillary force) are pinned by closed
 * forms of their geometry (tests/test_oracle_contact_angle.py).
 */
#ifndef CGL_ORACLE_H
#define CGL_ORACLE_H

#define OR_BC_PERIODIC 0
#define OR_BC_WALL 1
#define OR_BC_MIRROR 2

typedef struct or_state {
    int nx, ny, nz;
    int bc[6];               /* xlo, xhi, ylo, yhi, zlo, zhi */
    int npar;
    double sigma, rho0, nu_r, nu_b, beta;
    double fg;               /* repulsion scale Fg (S:84) */
    double rep_plane[6];     /* repulsion plane coordinate per face (S:1607-1612) */
    int rep_on[6];           /* 1 if that face repels (S:1633-1650) */
    double dom_lo[3], dom_hi[3]; /* Lagrange-point in-domain box (S:1832) */
    int dom_lo_on[3], dom_lo_strict[3], dom_hi_on[3];
    int ref_rules;           /* reference 1/8-symmetry rules in movement (S:1272-1335) */
    /* fields */
    double *ri, *bi, *fi, *ffi;                  /* [N][19] */
    double *rho, *rho_r, *rho_b, *phi, *ux, *uy, *uz; /* [N] */
    double *gx, *gy, *gz, *nxv, *nyv, *nzv;      /* [N] */
    int *flag;                                   /* [N] (LOGICAL as int) */
    int *walls, *walls_old, *walls_all, *pin, *pout; /* [N] */
    /* particles, [npar][3] unless noted */
    double *center, *center_old, *center0;
    double *up, *rp;
    double *fp, *tp, *frep, *tension, *tension_t;
    double *rad, *mass, *inertia, *cita, *shb;   /* [npar] */
} or_state;

void or_tables(double *wi, int *c, int *opc, int *sym, double *M, double *Minv);
void or_equilf_srt(double *feq, double rho, double ux, double uy, double uz);
void or_equilf(double *meq, double *fk, double rho, double ux, double uy, double uz,
               double fx, double fy, double fz);

void or_calmacros(or_state *s);
void or_derivatives(or_state *s);
void or_collision(or_state *s);
void or_recolor(or_state *s);
void or_boundary(or_state *s);
void or_streaming(or_state *s, double *f);
void or_curve(or_state *s);
void or_set_walls(or_state *s);
void or_refilling(or_state *s);
void or_momentum(or_state *s);
void or_lagrange_capillary(or_state *s);
void or_repulsive(or_state *s);
void or_movement(or_state *s);
void or_step(or_state *s, int particles);
void or_run(or_state *s, int nsteps, int particles);

#endif
