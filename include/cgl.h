/*
 * cgl.h -- C ABI of libcgl, the MI355X (gfx950) per-timestep lattice update of
 * the D3Q19 two-component colour-gradient MRT solver (X-Sheen/Symmetry,
 * Source1.f90, cited as S:N).
 *
 * The reference's time loop (S:215-252) calls external Fortran subroutines
 * with implicit interfaces and whole arrays passed by address.  libcgl takes
 * the place of that loop body for one GPU: a Fortran driver (or any C caller)
 * holds host arrays in the reference's column-major layout,
 *   populations (0:18, zl, yl, xl)  -> double[xl][yl][zl][19] in C terms,
 *   scalars     (zl, yl, xl)        -> double[xl][yl][zl],
 * uploads them once, advances the state on the device and downloads fields
 * only at output cadence.  All pointers are host pointers borrowed for the
 * duration of the call; device memory is owned by the handle.
 *
 * Conventions: every function returns CGL_OK (0) or a negative CGL_ERR_*
 * code; cgl_last_error() describes the last failure of the calling thread.
 * One host thread per handle; a handle is not thread-safe.  LOGICAL arrays
 * cross the ABI as int32 (0 / 1).
 */
#ifndef CGL_H
#define CGL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CGL_OK 0
#define CGL_ERR_ARG (-1)         /* invalid argument or size */
#define CGL_ERR_HIP (-2)         /* HIP runtime error (no GPU, out of memory, launch failure) */
#define CGL_ERR_UNSUPPORTED (-3) /* valid request the library does not implement */
#define CGL_ERR_RANGE (-4)       /* a density left the range the scatter form's fixed-point sums hold */

#define CGL_BC_PERIODIC 0 /* periodic face pair (the reference's next_* tables, S:295-324) */
#define CGL_BC_WALL 1     /* half-way bounce-back wall with a ghost layer (S:1557-1590) */
#define CGL_BC_MIRROR 2   /* specular mirror ghost layer (S:1443-1542) */

#define CGL_PREC_F64 0 /* FP64 storage and collision ("plain") */
#define CGL_PREC_F32 1 /* FP32 storage shifted by w_i*rho0*(1 +- phi)/2 of the reading node, FP32 collision */
#define CGL_PREC_F16 2 /* FP16 shifted storage, FP32 phi and collision */

#define CGL_LAYOUT_RB 0 /* red and blue populations, 2 x 19 structure-of-arrays */
#define CGL_LAYOUT_FS 1 /* factorised: total populations (19 SoA) + per-node recolour scalars,
                           r_i = A f_i + w_i G.c_i (the form Recolor produces, S:713-726); with
                           walls, mirrors or particles, colour side arrays for the slots of solid
                           and ghost nodes */

typedef struct cgl_handle cgl_t;

/* Physical and numerical parameters.  Names follow the reference's module
 * const / main (S:7-26, S:57-84).  Layout is fixed for Fortran bind(C). */
typedef struct cgl_params {
    int32_t nx, ny, nz;  /* xl, yl, zl (S:9-11) */
    int32_t bc[6];       /* faces x-lo, x-hi, y-lo, y-hi, z-lo, z-hi: CGL_BC_* */
    int32_t precision;   /* CGL_PREC_* */
    int32_t npar;        /* number of particles (par, S:24); 0 = fluid only */
    int32_t xchunk;      /* x planes per workgroup of the fused kernel; 0 = autotune on the first steps */
    int32_t device;      /* HIP device ordinal */
    int32_t layout;      /* CGL_LAYOUT_*: device storage of the populations */
    int32_t tile_rows;   /* rows per workgroup tile of the scatter-form step: 4, 8 or 16; 0 = 8, or autotuned with
                            xchunk = 0 (batches and the other kernels run their fixed tiles) */
    double sigma;        /* CSF surface-tension coefficient (sigma, S:66) */
    double rho0;         /* reference density (rho0, S:67) */
    double nu_r, nu_b;   /* kinematic viscosities of red and blue (S:69-70) */
    double beta;         /* recolouring parameter (beta, S:17) */
} cgl_params;

/* Allocate device state for one box.  Replaces the allocation block of the
 * reference's main (S:144-179).  Faces may be periodic (in pairs),
 * half-way bounce-back walls (S:1557-1590) or specular mirrors
 * (S:1443-1542), the last two with a ghost layer; npar > 0 adds moving
 * particles.  Walls, mirrors and particles run on both layouts. */
int cgl_create(cgl_t **h, const cgl_params *p);
int cgl_destroy(cgl_t *h);

/* Upload the post-streaming populations ri, bi (S:169-170 layout) as the
 * state at the top of a time step (what calmacros, S:216, reads). */
int cgl_upload_state(cgl_t *h, const double *ri, const double *bi);

/* Download the current state in the same layout: equal to the reference's
 * ri, bi after its streaming and curve calls (S:249-251) at live nodes; the
 * wall and mirror ghost layers are not part of the state and read as 0. */
int cgl_download_state(cgl_t *h, double *ri, double *bi);

/* Initial condition on the device: liquid cylinder along z through (xc, yc)
 * in 1-based lattice coordinates, R(z) = r0 + sum_m amp[m] cos(2 pi (m+1)
 * (z-1)/nz), phi = tanh(kprof (R(z) - r)), Laplace density jump 3 sigma/R(z)
 * inside, SRT equilibria at rest (the form of initialization, S:850-896). */
int cgl_init_cylinder(cgl_t *h, double xc, double yc, double r0, double kprof, int32_t nmodes, const double *amp);

/* Device initial state of a flat interface: red below z = z0 (1-based
 * lattice units), phi = tanh(kprof (z0 - z)), density rho0, at rest, with the
 * SRT equilibria of S:883-889 (the particle-at-interface workloads). */
int cgl_init_interface(cgl_t *h, double z0, double kprof);

/* Advance nsteps time steps: the loop body S:216-251 (calmacros,
 * derivatives, collision, Recolor, boundary, streaming) without output. */
int cgl_step(cgl_t *h, int32_t nsteps);

/* Moments of the current state (calmacros, S:770-831): rho, rho_R, rho_B,
 * bare velocity ux, uy, uz and phi (rho_N) in (zl, yl, xl) layout.  Any
 * pointer may be NULL.  On wall and mirror ghost layers phi is the copy of
 * the adjacent live layer (S:806-807) and the other fields are 0. */
int cgl_calmacros(cgl_t *h, double *rho, double *rho_r, double *rho_b, double *ux, double *uy, double *uz,
                  double *phi);

/* Phase-field gradient, unit normal and interface flag of the current state
 * (derivatives, S:498-533).  Any pointer may be NULL. */
int cgl_derivatives(cgl_t *h, double *gradx, double *grady, double *gradz, double *n_x, double *n_y, double *n_z,
                    int32_t *flag);

/* The fused kernel's x chunk in use and whether the autotuner (xchunk = 0)
 * is still timing candidates on the first steps.  No counterpart in the
 * reference. */
int cgl_tuned_xchunk(cgl_t *h, int32_t *xchunk, int32_t *tuning);

/* The launch shape the steps use, as the autotuner (xchunk = 0 at create)
 * chose it on the first steps, one parameter after the other with the
 * others held: the x chunk (x planes per workgroup, which is the number of
 * nodes each thread steps through along x: 2 .. 64), the tile order over
 * the XCDs (xcd_order: 0 launch order, 1 a contiguous range of tiles per
 * XCD, 2 whole z columns of tiles per XCD; fixed when CGL_XCD_REMAP is set)
 * and, for the scatter-form step, the tile height (tile_rows: 4, 8 or 16
 * rows of 64 nodes; fixed when cgl_params.tile_rows is).  Every choice
 * gives the same results bit for bit.  tuning = 1 while candidates are still
 * being timed.  Any pointer may be NULL.  No counterpart in the reference. */
int cgl_tuned_params(cgl_t *h, int32_t *xchunk, int32_t *xcd_order, int32_t *tile_rows, int32_t *tuning);

/* 1 when the handle's steps compute the next phase field in scatter form
 * (calmacros' sums S:792-799 accumulated by the writers of the populations),
 * so that a step reads the state once: FP32/FP16 factorised boxes, single
 * or batched, with ny % rows != 1 (rows = tile_rows, 8 when 0) and
 * nz % 64 != 1 (every workgroup tile has two rows and two lanes); single
 * boxes also with wall faces and particles (not mirror faces; rows 8, and
 * with particles from the second step after cgl_set_particles on), unless the
 * environment sets CGL_SCATTER=0 at creation.  0: a moments pass per step (gather form).
 * No counterpart in the reference. */
int cgl_phase_field_form(cgl_t *h, int32_t *scatter);

/* Checkpoint and restart (the reference has none).  cgl_checkpoint_save
 * writes the device state as the handle holds it (storage precision,
 * layout, both phase-field buffers, particle masks, band fields and particle
 * state) and the step bookkeeping to the file `path`; cgl_checkpoint_load
 * restores it into a handle created with the same parameters (box, faces,
 * precision, layout, npar or batch), after which the run continues bit for
 * bit as if it had not stopped.  The file is for this library on a host of
 * the same byte order, not a portable format.  A load that fails after the
 * header checks leaves the state undefined: upload or load again. */
int cgl_checkpoint_save(cgl_t *h, const char *path);
int cgl_checkpoint_load(cgl_t *h, const char *path);

/* Block until all queued device work of the handle has finished.
 * FP32/FP16 factorised boxes in the scatter form sum the next step's colour densities in
 * 32-bit fixed point at 2^29 units per rho0 (the scatter form): a colour
 * density of 4 rho0 or more does not fit.  Such a step raises a flag that
 * cgl_sync, cgl_time_steps and every call that reads the state (calmacros,
 * downloads, diagnostics, checkpoint save) report as CGL_ERR_RANGE; the
 * state is then lost.  Installing a state (upload, init, checkpoint load)
 * clears the flag; CGL_SCATTER=0 in the environment selects the gather form,
 * which has no such limit. */
int cgl_sync(cgl_t *h);

/* Particles (npar of the handle), in the reference's layout: per-particle
 * vectors are (npar, 3) Fortran arrays, component-major, like center(par,3)
 * (S:45); velocities and angular velocities pack upx/upy/upz and rpx/rpy/rpz
 * (S:49) the same way (NULL = at rest).  Centres are 1-based lattice
 * coordinates.  Then radius, mass, moment of inertia and contact angle
 * (radians) per particle, and the repulsion scale Fg (S:84).  Computes the
 * masks as the start-up set_walls does (S:212, S:326-395).  Replaces the
 * particle set-up of S:76-84 and initia_center (S:926-976). */
int cgl_set_particles(cgl_t *h, const double *center, const double *up, const double *rp, const double *rad,
                      const double *mass, const double *inertia, const double *cita, double fg);

/* Particle geometry that the reference hard-codes for its own box: the
 * repulsion planes of the wall force (S:1607-1612), rep_plane[6] / rep_on[6]
 * per face (x-lo, x-hi, y-lo, y-hi, z-lo, z-hi; 1-based lattice positions);
 * the box inside which Lagrange points count (S:1832), dom_lo[3],
 * dom_hi[3] and dom_on[9] = (low bound on, low bound strict, high bound on)
 * per axis (a point counts if x >= lo, or x > lo when strict, and x < hi);
 * and sym_rules = 1 for the symmetry-plane rules of movement (S:1272-1335),
 * which act on particles whose set-up centre lies on the plane x = xl - 1/2,
 * y = 3/2, y = yl - 1/2 or z = zl - 1/2 (the mirror planes of the reference's
 * 1/8 domain), and average the x and z components of particle par/3 + 1.
 * cgl_set_particles resets the defaults: planes half-way between the ghost
 * layer and the first live layer of every wall face, no box, no rules. */
int cgl_set_particle_geometry(cgl_t *h, const double *rep_plane, const int32_t *rep_on, const double *dom_lo,
                              const double *dom_hi, const int32_t *dom_on, int32_t sym_rules);

/* Particle state after the last step: centre, velocity, angular velocity,
 * momentum-exchange force and torque (Fpx.., Tpx.., S:1247-1252), capillary
 * force and torque (tension, tension_t, S:2020-2022), repulsion (Frex..,
 * S:1633-1665), each an (npar, 3) component-major array as above.  Any
 * pointer may be NULL. */
int cgl_get_particles(cgl_t *h, double *center, double *up, double *rp, double *force, double *torque,
                      double *fcap, double *tcap, double *frep);

/* Shb of each particle after the last step: the area of its sphere that
 * lies in the red fluid, the sum of the Lagrange points' surface elements
 * dS where phi > 0 (Lagrange_points, S:1864-1866), unscaled (output4 scales
 * it by 8 / (pi Rad_c^2), S:1793).  npar values. */
int cgl_get_particle_shb(cgl_t *h, double *shb);

/* Masks of set_walls in force (walls, walls_all, pin, pout; S:326-395),
 * (zl, yl, xl) layout; within a staged step after cgl_set_walls, the new
 * masks.  Any pointer may be NULL. */
int cgl_download_masks(cgl_t *h, int32_t *walls, int32_t *walls_all, int32_t *pin, int32_t *pout);

/* A time step with particles, stage by stage: the reference calls these
 * routines one after another in its time loop (S:218-251), and each entry
 * point runs exactly the kernels cgl_step runs at that point, so a staged
 * step equals a cgl_step step bit for bit.  A staged step opens with
 * cgl_calmacros (calmacros, S:770-831; any output pointer may be NULL) and
 * continues in this order; an entry point called out of order returns
 * CGL_ERR_ARG, and while a staged step is open cgl_step, cgl_time_steps and
 * the calls that recompute moments (calmacros, derivatives, diagnostics,
 * downloads, checkpoint save) refuse.  cgl_get_particles,
 * cgl_get_particle_shb, cgl_download_masks and cgl_particle_band read the
 * state between stages.  Installing a state or particles abandons an open
 * staged step.  Wall-only handles (npar = 0) go from cgl_calmacros straight
 * to cgl_collide_stream. */
int cgl_capillary(cgl_t *h);      /* Lagrange_points + capillary_force2 (S:1799-2028): fcap, tcap, Shb */
int cgl_momentum(cgl_t *h);       /* momentum (S:1178-1261): force and torque of the links */
int cgl_repulsive(cgl_t *h);      /* repulsive (S:1594-1671): wall and pair repulsion */
int cgl_move(cgl_t *h);           /* movement (S:1263-1349): symmetry rules, velocities, centres */
int cgl_set_walls(cgl_t *h);      /* set_walls (S:326-395): masks of the new centres */
int cgl_refill(cgl_t *h);         /* refilling (S:1351-1418): nodes the move uncovered */
/* derivatives (S:452-597), collision (S:599-695), Recolor (S:697-732),
 * boundary (S:1544-1592), streaming of r and b (S:735-768) and curve
 * (S:990-1034): one fused pull kernel plus the bounce-back links (the pull
 * layout streams as it reads, so the reference's collision, Recolor,
 * boundary and streaming calls have no separate entry points).  Ends the
 * staged step. */
int cgl_collide_stream(cgl_t *h);

/* The FP64 macroscopic fields the particle stages read and write in the band
 * around the particles (rho, rho_R, rho_B, ux, uy, uz in (zl, yl, xl)
 * layout): calmacros' values there, and refilling's at the nodes it
 * refilled (S:1364-1391).  Defined at the band's nodes only (within R + 3 of
 * a centre); any pointer may be NULL. */
int cgl_particle_band(cgl_t *h, double *rho, double *rho_r, double *rho_b, double *ux, double *uy, double *uz);

/* Global FP64 sums on the device over the live fluid nodes (ghost layers and
 * particle interiors excluded) of the current state's macroscopic fields:
 * sums = {mass, red mass, blue mass, momentum x, y, z, kinetic energy}
 * (the mass check and Energy.dat sums of S:1673-1797).  With slice_red,
 * also the red area sum (1 + phi)/2 of every z slice (nz values), from which
 * the neck radius of a jet along z is sqrt(area / pi).  Sums run in a fixed
 * order, so they are reproducible. */
int cgl_diagnostics(cgl_t *h, double *sums, double *slice_red);

/* cgl_diagnostics without any device allocation: the same sums and slice
 * areas (equal to rounding: the sums run in another fixed order), formed by
 * kernels that read the state and write their partial sums to a static
 * device buffer of CGL_NOALLOC_SCRATCH_DOUBLES doubles (4 MiB per GPU, held
 * until the process ends); no hipMalloc or other allocating call.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * (code under HIP graph capture, serving loops, harnesses that forbid it).
 * A box needs 8 + nz + 7 gz gy + gy nz doubles (gz = ceil(nz/64),
 * gy = ceil(ny/4)): larger boxes get CGL_ERR_UNSUPPORTED.  Every call on one
 * GPU shares that GPU's buffer, so calls that overlap on one GPU (from other
 * threads or streams) race; a call returns after its results are copied. */
#define CGL_NOALLOC_SCRATCH_DOUBLES (1 << 19)
int cgl_diagnostics_noalloc(cgl_t *h, double *sums, double *slice_red);

/* ---------------------------------------------------------------- batches
 * Many independent periodic boxes of uneven sizes advanced together: one
 * launch per kernel covers the tiles of every box (nodes packed box after
 * box), and cgl_step replays the steps from a captured HIP graph.  The
 * reference runs one box per process; there is no counterpart call. */
typedef struct cgl_box {
    int32_t nx, ny, nz; /* box size, as xl, yl, zl */
    int32_t reserved;   /* must be 0 */
    double nu_r, nu_b;  /* kinematic viscosities of the two fluids (tau per box) */
} cgl_box;

/* precision, layout, sigma, rho0, beta, xchunk and device come from `p`
 * (its nx, ny, nz, nu_r, nu_b are ignored); faces must be periodic and npar
 * zero.  cgl_step, cgl_sync, cgl_time_steps and cgl_destroy take the batch
 * handle; per-box data go through the cgl_box_* calls. */
int cgl_batch_create(cgl_t **h, const cgl_params *p, int32_t nbox, const cgl_box *boxes);
/* As cgl_upload_state / cgl_init_cylinder for box `box`; every box is set
 * before the batch's first step. */
int cgl_box_upload_state(cgl_t *h, int32_t box, const double *ri, const double *bi);
int cgl_box_init_cylinder(cgl_t *h, int32_t box, double xc, double yc, double r0, double kprof, int32_t nmodes,
                          const double *amp);
/* As cgl_download_state / cgl_calmacros for box `box`. */
int cgl_box_download_state(cgl_t *h, int32_t box, double *ri, double *bi);
int cgl_box_calmacros(cgl_t *h, int32_t box, double *rho, double *rho_r, double *rho_b, double *ux, double *uy,
                      double *uz, double *phi);

/* Time nsteps steps on the device with HIP events recorded on the handle's
 * stream: total elapsed ms, and ms summed over the moments (scatter form: the
 * phase-field pass) and fused-collide kernels; with walls or particles, the
 * moments and collide kernels alone (in the scatter form the finishing pass,
 * and the tile and band step kernels; the particle kernels between them are
 * in neither).  No counterpart in the reference (it has no timers). */
int cgl_time_steps(cgl_t *h, int32_t nsteps, double *ms_total, double *ms_moments, double *ms_collide);

/* Achievable HBM bandwidth of a device: a streaming copy of `bytes` bytes,
 * repeated reps times, read + write bytes per second in GB/s.  Also the
 * known-size kernel used to calibrate rocprofv3 byte counters.  No
 * counterpart in the reference. */
int cgl_probe_copy_bandwidth(int32_t device, int64_t bytes, int32_t reps, double *gbps);

/* Device name of the handle (for logs). */
int cgl_device_name(cgl_t *h, char *buf, int32_t len);

const char *cgl_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* CGL_H */
