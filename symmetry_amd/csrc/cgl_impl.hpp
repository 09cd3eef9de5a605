// cgl_impl.hpp -- device handle and per-precision kernel launchers.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "../../include/cgl.h"
#include "cgl_kernels.hpp"
#include "cgl_particles.hpp"
#include "cgl_scatter.hpp"

struct cgl_handle {
    cgl_params p;
    cgl::Lat L;
    size_t N;
    size_t esize;     // bytes per stored population value
    size_t phisize;   // bytes per phi value
    int layout;       // CGL_LAYOUT_* of the stepping state
    int state_layout; // layout the current state is held in (RB after upload / init)
    void *pops[2][2]; // RB: [buffer][colour], colour 0 red, 1 blue, 19*N values each.
                      // FS layout: only pops[0] (staging for upload / init).
    void *fsf[2];     // FS: total populations, 19*N values per buffer
    void *fss[2];     // FS: recolour scalars, NS*N values per buffer
    void *phi[2];
    int cur;          // buffer index holding the current state
    int phicur;       // phi buffer holding the shift reference of the current state
    int xc;           // x planes per workgroup of the fused kernel
    int tuning;       // 1 while the first steps still time candidate x chunks
    int tune_idx;     // next candidate
    int tune_best_val;   // best value of the dimension being tuned so far
    float tune_best_ms;  // its time, < 0 before the dimension's first candidate
    double *mac;      // scratch: 7*N doubles for macroscopic downloads (lazy)
    double *aos;      // scratch: 19*N doubles for host-layout conversion (lazy)
    int *iflag;       // scratch: N ints (lazy)
    size_t scratch_n; // nodes the scratch buffers are sized for
    hipStream_t stream;
    hipEvent_t ev[4];
    int steps_done;
    // walls and particles
    int ext;             // 1: wall faces or particles present (EXT kernels)
    int npar;
    int primed;          // start-up derivatives done for the current state
    int cold;            // first step after set_particles: ffi = fi (S:894)
    cgl::Ext E;          // wall faces and device pointers (ntype set per launch)
    unsigned *ntype[2];  // node-type masks, double buffered (the previous step's are walls_old)
    int ntcur;           // buffer holding the masks in force
    int *walls_i, *pin_i, *pout_i;
    cgl::Part P;         // particle state (device pointers)
    // batches of independent boxes (nbox > 0): packed node index
    int nbox;
    cgl::BoxDev *dboxes; // device table
    cgl::BoxDev *hboxes; // host copy
    unsigned tiles1, tiles2;
    // steady-state steps replayed from a captured graph (two steps per replay)
    int use_graph;       // replay steady-state steps from a captured graph
    hipGraphExec_t gexec;
    int gkey;            // cur | phicur << 1 | state_layout << 2 | ntcur << 3 | phi_ok << 4 at capture, -1 = none
    // periodic FP32/FP16 factorised boxes: next phase field in scatter form
    int sc;              // 1: steady steps use k_collide_sc + k_phi_sc
    int remap_env;       // 1: CGL_XCD_REMAP was set, and its value holds for every kernel
    int sc_remap;        // tile order of the scatter step (tuned; CGL_XCD_REMAP when set)
    int ky;              // rows per workgroup tile of the scatter step: 4, 8 or 16 (tuned)
    int tile_rows;       // cgl_params.tile_rows: 0 = tuned, else the fixed ky
    int tune_dim;        // dimension being tuned: 0 x chunk, 1 tile order, 2 tile rows
    int phi_ok;          // phi[phicur ^ 1] holds the current state's phase field (from the last step)
    int scal_stale;      // the current state's recolour scalars were not written (mid-run only)
    cgl::ScBufs sb;      // fixed-point partial sums
    hipEvent_t *tev;     // timing runs: events around the wall/particle step's moments and collide kernels
    int stage;           // staged step: the last stage run (STAGE_*), STAGE_NONE between steps
    unsigned band_blocks; // workgroups per particle of the band kernels (one thread per node of its band box)
    unsigned nrm_blocks;  // workgroups per particle of k_band_normals (its box with margin R + 4)
    float *bnrm;          // scatter form with particles: the band boxes' unit normals (3 x N)
};

// Stages of a step driven through the per-stage entry points, in order.
enum { STAGE_NONE = 0, STAGE_CALMACROS, STAGE_CAPILLARY, STAGE_MOMENTUM, STAGE_REPULSIVE, STAGE_MOVE, STAGE_SET_WALLS,
       STAGE_REFILL };

namespace cgl {

inline dim3 node_grid(const Lat &L) { return dim3((L.nz + TZ - 1) / TZ, (L.ny + TY - 1) / TY, L.nx); }
inline dim3 node_block() { return dim3(TZ, TY, 1); }
template <class R> inline dim3 k2_block() { return dim3(TZ, KTY<R>, 1); }
template <class R> inline dim3 k2_grid(const Lat &L, int xc)
{
    return dim3((L.nz + TZ - 1) / TZ, (L.ny + KTY<R> - 1) / KTY<R>, (L.nx + xc - 1) / xc);
}

// Input buffers of the current state and output buffers of the next one.
inline void state_in(const cgl_handle *h, const void *&i0, const void *&i1)
{
    if (h->state_layout == RB) {
        i0 = h->pops[h->cur][0];
        i1 = h->pops[h->cur][1];
    } else {
        i0 = h->fsf[h->cur];
        i1 = h->fss[h->cur];
    }
}
inline void state_out(const cgl_handle *h, void *&o0, void *&o1)
{
    if (h->layout == RB) {
        o0 = h->pops[h->cur ^ 1][0];
        o1 = h->pops[h->cur ^ 1][1];
    } else {
        o0 = h->fsf[h->cur ^ 1];
        o1 = h->fss[h->cur ^ 1];
    }
}

// x planes per workgroup of the x-marching moments kernel (factorised layout)
constexpr int K1_XC = 8;

template <class S, class R, class P, bool SHIFT> struct Impl {
    // phi buffer that moments() writes: the step's phi buffer, or the spare one
    // while that already holds the scatter step's phase field (observing the
    // state does not change the next step)
    static int phi_mom(const cgl_handle *h) { return h->phi_ok ? h->phicur : h->phicur ^ 1; }
    // enc: the scatter form's phase-field buffer (signed minority fractions)
    static void moments(cgl_handle *h, bool full, bool enc = false)
    {
        const void *i0, *i1;
        state_in(h, i0, i1);
        const P *phio = (const P *)h->phi[h->phicur];
        P *phi = (P *)h->phi[phi_mom(h)];
        double *mac = full ? h->mac : nullptr;
        auto launch = [&](auto lin) {
            constexpr int LIN = decltype(lin)::value;
            if (full)
                k_moments<LIN, S, R, P, SHIFT, true>
                    <<<node_grid(h->L), node_block(), 0, h->stream>>>(h->L, i0, i1, phio, phi, mac);
            else
                k_moments<LIN, S, R, P, SHIFT, false>
                    <<<node_grid(h->L), node_block(), 0, h->stream>>>(h->L, i0, i1, phio, phi, mac);
        };
        if (h->state_layout == RB) {
            launch(std::integral_constant<int, RB>{});
        } else if (!full) {
            const int xc = K1_XC;
            dim3 grid((h->L.nz + TZ - 1) / TZ, (h->L.ny + TY - 1) / TY, (h->L.nx + xc - 1) / xc);
            if constexpr (SHIFT && std::is_same_v<P, float>) {
                if (enc) {
                    k_moments_fsx<S, R, P, SHIFT, true><<<grid, node_block(), 0, h->stream>>>(h->L, i0, i1, phio, phi, xc);
                    return;
                }
            }
            k_moments_fsx<S, R, P, SHIFT><<<grid, node_block(), 0, h->stream>>>(h->L, i0, i1, phio, phi, xc);
        } else {
            launch(std::integral_constant<int, FS>{});
        }
    }
    // one fused step; afterwards the state is held in h->layout
    static void collide(cgl_handle *h)
    {
        const void *i0, *i1;
        void *o0, *o1;
        state_in(h, i0, i1);
        state_out(h, o0, o1);
        const P *phio = (const P *)h->phi[h->phicur];
        const P *phi = (const P *)h->phi[h->phicur ^ 1];
        const dim3 grid = k2_grid<R>(h->L, h->xc);
        if (h->state_layout == RB && h->layout == RB)
            k_collide<RB, RB, S, R, P, SHIFT>
                <<<grid, k2_block<R>(), 0, h->stream>>>(h->L, i0, i1, o0, o1, phi, phio, h->xc);
        else if (h->state_layout == RB)
            k_collide<RB, FS, S, R, P, SHIFT>
                <<<grid, k2_block<R>(), 0, h->stream>>>(h->L, i0, i1, o0, o1, phi, phio, h->xc);
        else
            k_collide<FS, FS, S, R, P, SHIFT>
                <<<grid, k2_block<R>(), 0, h->stream>>>(h->L, i0, i1, o0, o1, phi, phio, h->xc);
        h->state_layout = h->layout;
    }
    // ---- batches: one launch per kernel over the tiles of every box
    static void batch_moments(cgl_handle *h, bool enc = false)
    {
        const void *i0, *i1;
        state_in(h, i0, i1);
        const P *phio = (const P *)h->phi[h->phicur];
        P *phi = (P *)h->phi[h->phicur ^ 1];
        const int remap = h->L.xcd_remap;
        if (h->state_layout == RB)
            k_moments_batch<RB, S, R, P, SHIFT>
                <<<h->tiles1, node_block(), 0, h->stream>>>(h->dboxes, h->nbox, remap, K1_XC, i0, i1, phio, phi);
        else {
            if constexpr (SHIFT && std::is_same_v<P, float>) {
                if (enc) {
                    k_moments_batch<FS, S, R, P, SHIFT, true><<<h->tiles1, node_block(), 0, h->stream>>>(
                        h->dboxes, h->nbox, remap, K1_XC, i0, i1, phio, phi);
                    return;
                }
            }
            k_moments_batch<FS, S, R, P, SHIFT>
                <<<h->tiles1, node_block(), 0, h->stream>>>(h->dboxes, h->nbox, remap, K1_XC, i0, i1, phio, phi);
        }
    }
    static void batch_collide(cgl_handle *h)
    {
        const void *i0, *i1;
        void *o0, *o1;
        state_in(h, i0, i1);
        state_out(h, o0, o1);
        const P *phio = (const P *)h->phi[h->phicur];
        const P *phi = (const P *)h->phi[h->phicur ^ 1];
        const int remap = h->L.xcd_remap;
        if (h->state_layout == RB && h->layout == RB)
            k_collide_batch<RB, RB, S, R, P, SHIFT><<<h->tiles2, k2_block<R>(), 0, h->stream>>>(
                h->dboxes, h->nbox, remap, h->xc, i0, i1, o0, o1, phi, phio);
        else if (h->state_layout == RB)
            k_collide_batch<RB, FS, S, R, P, SHIFT><<<h->tiles2, k2_block<R>(), 0, h->stream>>>(
                h->dboxes, h->nbox, remap, h->xc, i0, i1, o0, o1, phi, phio);
        else
            k_collide_batch<FS, FS, S, R, P, SHIFT><<<h->tiles2, k2_block<R>(), 0, h->stream>>>(
                h->dboxes, h->nbox, remap, h->xc, i0, i1, o0, o1, phi, phio);
        h->state_layout = h->layout;
    }
    // derivatives of the phi written by the last moments() call
    static void derivatives(cgl_handle *h, double *out, int *flag)
    {
        k_derivatives<P><<<node_grid(h->L), node_block(), 0, h->stream>>>(h->L, (const P *)h->phi[phi_mom(h)], out,
                                                                         flag);
    }
    // ---- scatter-form step (periodic FP32/FP16 factorised boxes)
    // the current state's phase field in gather form, when no scatter step
    // has produced it
    static void sc_pre(cgl_handle *h)
    {
        if (h->phi_ok) return;
        if (h->nbox)
            batch_moments(h, true);
        else
            moments(h, false, true);
    }
    static void sc_collide(cgl_handle *h, bool last)
    {
        if constexpr (SHIFT && std::is_same_v<P, float>) {
            const S *fin = (const S *)h->fsf[h->cur];
            S *fout = (S *)h->fsf[h->cur ^ 1];
            float4 *sout = last ? (float4 *)h->fss[h->cur ^ 1] : nullptr;
            const float *phi = (const float *)h->phi[h->phicur ^ 1];
            // tile order h->sc_remap (tuned; by default whole z columns of
            // tiles on one XCD, consecutive columns over the XCDs: at
            // 256x256x512 +2 % for FP32 and +1.4 % for FP16 storage over the
            // XCD-contiguous order) and tile height h->ky (tuned, 8 rows by
            // default); the FP32 batch takes launch order and 8 rows
            Lat Ls = h->L;
            Ls.xcd_remap = h->sc_remap;
            if (h->nbox) {
                k_collide_sc_batch<S><<<h->tiles2, k2_block<R>(), 0, h->stream>>>(
                    h->dboxes, h->nbox, sizeof(S) == 4 && !h->remap_env ? 0 : h->L.xcd_remap, h->xc, fin, fout,
                    sout, phi, h->sb);
            } else {
                const bool one = (double)Q * (double)h->N * sizeof(S) < 4294967296.0; // one base for all directions
                auto launch = [&](auto rows) {
                    constexpr int KY = decltype(rows)::value;
                    const dim3 grid((h->L.nz + TZ - 1) / TZ, (h->L.ny + KY - 1) / KY, (h->L.nx + h->xc - 1) / h->xc);
                    if (one)
                        k_collide_sc<S, true, KY>
                            <<<grid, dim3(TZ, KY, 1), 0, h->stream>>>(Ls, fin, fout, sout, phi, h->sb, h->xc);
                    else
                        k_collide_sc<S, false, KY>
                            <<<grid, dim3(TZ, KY, 1), 0, h->stream>>>(Ls, fin, fout, sout, phi, h->sb, h->xc);
                };
                if (h->ky == 4)
                    launch(std::integral_constant<int, 4>{});
                else if (h->ky == 16)
                    launch(std::integral_constant<int, 16>{});
                else
                    launch(std::integral_constant<int, 8>{});
            }
            h->scal_stale = !last;
        }
    }
    // the next step's phase field into the spare phi buffer (it becomes the
    // step's buffer when the parities flip)
    static void sc_fin(cgl_handle *h)
    {
        if (h->nbox)
            k_phi_sc_batch<<<h->tiles1, node_block(), 0, h->stream>>>(h->dboxes, h->nbox, K1_XC, h->xc, h->sb,
                                                                       (float *)h->phi[h->phicur]);
        else if (h->L.nz % 2 == 0 && h->L.nx % 2 == 0) // four nodes per thread
            k_phi_sc4<<<dim3((h->L.nz / 2 + TZ - 1) / TZ, (h->L.ny + TY - 1) / TY, h->L.nx / 2), node_block(), 0,
                        h->stream>>>(h->L, h->sb, h->xc, h->ky, (float *)h->phi[h->phicur], h->ext ? SC_BIAS : 0);
        else if (h->L.nz % 2 == 0) // two nodes per thread
            k_phi_sc2<<<dim3((h->L.nz / 2 + TZ - 1) / TZ, (h->L.ny + TY - 1) / TY, h->L.nx), node_block(), 0,
                        h->stream>>>(h->L, h->sb, h->xc, h->ky, (float *)h->phi[h->phicur], h->ext ? SC_BIAS : 0);
        else
            k_phi_sc<<<node_grid(h->L), node_block(), 0, h->stream>>>(h->L, h->sb, h->xc, h->ky,
                                                                      (float *)h->phi[h->phicur],
                                                                      h->ext ? SC_BIAS : 0);
        h->phi_ok = 1;
    }
    static void step_sc(cgl_handle *h, bool last)
    {
        sc_pre(h);
        sc_collide(h, last);
        sc_fin(h);
    }
    static void phi_aos(cgl_handle *h, const double *ra, const double *ba)
    {
        unsigned nb = (unsigned)((h->N + 255) / 256);
        k_phi_aos<P><<<nb, 256, 0, h->stream>>>(h->L, ra, ba, (P *)h->phi[h->phicur]);
    }
    static void pack(cgl_handle *h, const double *aos, int red)
    {
        k_pack<S, P, SHIFT><<<node_grid(h->L), node_block(), 0, h->stream>>>(
            h->L, aos, (S *)h->pops[h->cur][red ? 0 : 1], (const P *)h->phi[h->phicur], red, h->E);
    }
    static void unpack(cgl_handle *h, double *aos, int red)
    {
        if (h->state_layout == RB)
            k_unpack<S, P, SHIFT><<<node_grid(h->L), node_block(), 0, h->stream>>>(
                h->L, (const S *)h->pops[h->cur][red ? 0 : 1], aos, (const P *)h->phi[h->phicur], red, h->E);
        else if (h->ext)
            k_unpack_fs_ext<S, P, SHIFT><<<node_grid(h->L), node_block(), 0, h->stream>>>(
                h->L, h->E, (const S *)h->fsf[h->cur], h->fss[h->cur], h->pops[h->cur][0], h->pops[h->cur][1], aos,
                red);
        else
            k_unpack_fs<S, SHIFT><<<node_grid(h->L), node_block(), 0, h->stream>>>(
                h->L, (const S *)h->fsf[h->cur], h->fss[h->cur], aos, red);
    }
    // ---- walls and particles
    static Ext ext_with(const cgl_handle *h, const unsigned *nt)
    {
        Ext E = h->E;
        E.ntype = h->npar ? nt : nullptr;
        // factorised layout: the side arrays of the current and the next state
        E.rbi[0] = h->pops[h->cur][0];
        E.rbi[1] = h->pops[h->cur][1];
        E.rbo[0] = h->pops[h->cur ^ 1][0];
        E.rbo[1] = h->pops[h->cur ^ 1][1];
        return E;
    }
    static dim3 pgrid(const cgl_handle *h, int slices) { return dim3(h->npar, slices, 1); }
    static bool ghost_faces(const cgl_handle *h)
    {
        for (int a = 0; a < 3; ++a)
            if (h->E.wall[a][0] || h->E.wall[a][1] || h->E.mirror[a][0] || h->E.mirror[a][1]) return true;
        return false;
    }
    static bool mirrors(const cgl_handle *h)
    {
        for (int a = 0; a < 3; ++a)
            if (h->E.mirror[a][0] || h->E.mirror[a][1]) return true;
        return false;
    }
    // launch shape of the face kernels: the largest face, six faces
    static dim3 face_grid(const cgl_handle *h)
    {
        int m = h->L.nx * h->L.ny;
        if (h->L.ny * h->L.nz > m) m = h->L.ny * h->L.nz;
        if (h->L.nx * h->L.nz > m) m = h->L.nx * h->L.nz;
        return dim3((m + 255) / 256, 6, 1);
    }
    // mirror fills of phi, and of the band's macroscopic fields, with particles
    static void mirror_phi(cgl_handle *h, P *phi)
    {
        if (mirrors(h)) k_wall_phi<P><<<face_grid(h), 256, 0, h->stream>>>(h->L, h->E, phi, 1);
    }
    static void mirror_mac(cgl_handle *h)
    {
        if (h->npar && mirrors(h)) k_mirror_mac<<<face_grid(h), 256, 0, h->stream>>>(h->L, h->E, h->E.mac);
    }
    // calmacros with walls and particles into the step's phi buffer, whose
    // entries are PP (P, or QPhi for the scatter form's buffer)
    static void moments_ext(cgl_handle *h, bool full)
    {
        moments_ext_t<P>(h, full);
        h->phi_ok = 0; // the buffer no longer holds a scatter step's phase field
    }
    template <class PP> static void moments_ext_t(cgl_handle *h, bool full)
    {
        const void *i0, *i1;
        state_in(h, i0, i1);
        const PP *phio = (const PP *)h->phi[h->phicur];
        PP *phi = (PP *)h->phi[h->phicur ^ 1];
        const Ext E = ext_with(h, h->ntype[h->ntcur]);
        if (h->state_layout == FS) {
            if (full) {
                if constexpr (std::is_same_v<PP, P>)
                    k_moments<FS, S, R, P, SHIFT, true, true>
                        <<<node_grid(h->L), node_block(), 0, h->stream>>>(h->L, i0, i1, phio, phi, h->mac, E);
            } else {
                const int xc = K1_XC;
                const dim3 grid((h->L.nz + TZ - 1) / TZ, (h->L.ny + TY - 1) / TY, (h->L.nx + xc - 1) / xc);
                k_moments_fsx_ext<S, R, PP, SHIFT><<<grid, node_block(), 0, h->stream>>>(h->L, i0, i1, phio, phi, xc, E);
                if (mirrors(h))
                    k_moments_special<S, R, PP, SHIFT, false>
                        <<<face_grid(h), 256, 0, h->stream>>>(h->L, i0, i1, phio, phi, E, h->P);
                if (h->npar)
                    k_moments_special<S, R, PP, SHIFT, true><<<dim3(h->band_blocks, h->npar), PB, 0, h->stream>>>(
                        h->L, i0, i1, phio, phi, E, h->P);
            }
        } else if constexpr (!std::is_same_v<PP, P>) {
            return; // the scatter form starts from a factorised state
        } else if (full) {
            k_moments<RB, S, R, P, SHIFT, true, true>
                <<<node_grid(h->L), node_block(), 0, h->stream>>>(h->L, i0, i1, phio, phi, h->mac, E);
        } else {
            k_moments<RB, S, R, P, SHIFT, false, true>
                <<<node_grid(h->L), node_block(), 0, h->stream>>>(h->L, i0, i1, phio, phi, nullptr, E);
        }
        wall_phi(h, phi);
        if (full) wall_phi(h, h->mac + 6 * h->N); // the FP64 phi output
        mirror_mac(h);
    }
    template <class T> static void wall_phi(cgl_handle *h, T *phi)
    {
        k_wall_phi<T><<<face_grid(h), 256, 0, h->stream>>>(h->L, h->E, phi);
    }
    // start-up derivatives (S:214) for the capillary force of the first step
    static void prime(cgl_handle *h)
    {
        moments_ext(h, false);
        P *phi = (P *)h->phi[h->phicur ^ 1];
        const Ext E = ext_with(h, h->ntype[h->ntcur]);
        k_ghost_phi<P><<<pgrid(h, LSLICES), PB, 0, h->stream>>>(h->L, E, h->P, h->ntype[h->ntcur], h->pin_i, phi);
        mirror_phi(h, phi);
        k_derivs_ext<P><<<node_grid(h->L), node_block(), 0, h->stream>>>(h->L, E, phi, nullptr, nullptr);
    }
    // diagnostics sums without a field buffer, into the caller's scratch:
    // scr[0..NDIAG) the sums, scr[8..8+nz) the slice areas, then the partials
    // (diag_scratch_doubles() of them in all)
    static size_t diag_scratch_doubles(const cgl_handle *h)
    {
        const size_t gz = (h->L.nz + TZ - 1) / TZ, gy = (h->L.ny + TY - 1) / TY;
        return 8 + (size_t)h->L.nz + NDIAG * gz * gy + gy * (size_t)h->L.nz;
    }
    static void diag_fused(cgl_handle *h, double *scr)
    {
        const void *i0, *i1;
        state_in(h, i0, i1);
        const P *phio = (const P *)h->phi[h->phicur];
        const dim3 grid((h->L.nz + TZ - 1) / TZ, (h->L.ny + TY - 1) / TY, 1);
        double *out = scr, *area = scr + 8, *part = area + h->L.nz, *spart = part + (size_t)NDIAG * grid.x * grid.y;
        const Ext E = ext_with(h, h->npar ? h->ntype[h->ntcur] : nullptr);
        if (h->ext && h->state_layout == FS)
            k_diag_tiles<FS, S, R, P, SHIFT, true>
                <<<grid, node_block(), 0, h->stream>>>(h->L, i0, i1, phio, E, part, spart);
        else if (h->ext)
            k_diag_tiles<RB, S, R, P, SHIFT, true>
                <<<grid, node_block(), 0, h->stream>>>(h->L, i0, i1, phio, E, part, spart);
        else if (h->state_layout == RB)
            k_diag_tiles<RB, S, R, P, SHIFT, false>
                <<<grid, node_block(), 0, h->stream>>>(h->L, i0, i1, phio, E, part, spart);
        else
            k_diag_tiles<FS, S, R, P, SHIFT, false>
                <<<grid, node_block(), 0, h->stream>>>(h->L, i0, i1, phio, E, part, spart);
        k_diag_final<<<1, 256, 0, h->stream>>>(part, grid.x * grid.y, spart, grid.y, h->L.nz, out, area);
    }
    static void derivatives_ext(cgl_handle *h, double *out, int *flag)
    {
        moments_ext(h, false);
        P *phi = (P *)h->phi[h->phicur ^ 1];
        const Ext E = ext_with(h, h->ntype[h->ntcur]);
        if (h->npar) {
            k_ghost_phi<P><<<pgrid(h, LSLICES), PB, 0, h->stream>>>(h->L, E, h->P, h->ntype[h->ntcur], h->pin_i, phi);
            mirror_phi(h, phi);
        }
        Ext E2 = E;
        E2.der = nullptr;
        k_derivs_ext<P><<<node_grid(h->L), node_block(), 0, h->stream>>>(h->L, E2, phi, out, flag);
    }
    // masks of the current particle centres into buffer `buf` (set_walls, S:326-395)
    static void build_masks(cgl_handle *h, int buf)
    {
        const Ext E = ext_with(h, h->ntype[buf]);
        k_masks_solid<<<pgrid(h, PSLICES), PB, 0, h->stream>>>(h->L, E, h->P, h->ntype[buf], h->walls_i);
        k_masks_links<<<pgrid(h, PSLICES), PB, 0, h->stream>>>(h->L, E, h->P, h->ntype[buf], h->walls_i, h->pin_i,
                                                               h->pout_i);
        if (mirrors(h))
            k_mirror_masks<<<face_grid(h), 256, 0, h->stream>>>(h->L, h->E, h->ntype[buf], h->walls_i, h->pin_i,
                                                                h->pout_i);
        build_lists(h, buf);
    }
    // the particles' link lists from the masks in buffer `buf`
    static void build_lists(cgl_handle *h, int buf)
    {
        const Ext E = ext_with(h, h->ntype[buf]);
        if (h->npar <= LISTS_WIDE_MAX)
            k_lists<1024><<<pgrid(h, LSLICES), 1024, 0, h->stream>>>(h->L, E, h->P, h->ntype[buf], h->walls_i,
                                                                     h->pin_i, h->pout_i);
        else
            k_lists<PB><<<pgrid(h, LSLICES), PB, 0, h->stream>>>(h->L, E, h->P, h->ntype[buf], h->walls_i, h->pin_i,
                                                                 h->pout_i);
    }
    // The stages of a step with walls and/or particles (S:216-251).  The
    // masks of the step's start stay in ntype[ntcur] until the step ends;
    // set_walls builds the new ones into ntype[ntcur ^ 1].
    static void st_calmacros(cgl_handle *h, bool full)
    {
        if (h->npar && !h->primed) {
            prime(h);
            h->primed = 1;
        }
        moments_ext(h, full); // calmacros with the masks in force
    }
    static void st_lagrange(cgl_handle *h)
    {
        const Ext Eo = ext_with(h, h->ntype[h->ntcur]);
        k_lagrange<P><<<pgrid(h, LAG_BLOCKS), PB, 0, h->stream>>>(h->L, Eo, h->P, (P *)h->phi[h->phicur ^ 1],
                                                                  h->walls_i);
    }
    static void st_momentum(cgl_handle *h)
    {
        const Ext Eo = ext_with(h, h->ntype[h->ntcur]);
        const S *rs = (const S *)h->pops[h->cur][0], *bs = (const S *)h->pops[h->cur][1];
        if (h->state_layout == FS)
            k_momentum_fs<S, SHIFT><<<pgrid(h, LSLICES), PB, 0, h->stream>>>(h->L, Eo, h->P,
                                                                          (const S *)h->fsf[h->cur], h->cold);
        else
            k_momentum<S, SHIFT><<<pgrid(h, LSLICES), PB, 0, h->stream>>>(h->L, Eo, h->P, h->ntype[h->ntcur],
                                                                          h->walls_i, h->pout_i, rs, bs, h->cold);
        h->cold = 0;
    }
    static void st_move(cgl_handle *h, int phases)
    {
        const Ext E = ext_with(h, h->ntype[h->ntcur]);
        if (phases & (MV_MOM | MV_CAP | MV_REP)) k_part_sums<<<h->npar, PB, 0, h->stream>>>(h->L, E, h->P, phases);
        if (phases & MV_MOVE) k_move<<<1, PB, 0, h->stream>>>(h->L, E, h->P, phases);
    }
    static void st_set_walls(cgl_handle *h)
    {
        // clear the spare buffer (built two steps ago) and the int masks
        // (built last step), then the new masks
        const Ext Eo = ext_with(h, h->ntype[h->ntcur]);
        const int nb = h->ntcur ^ 1;
        k_masks_clear<<<pgrid(h, PSLICES), PB, 0, h->stream>>>(h->L, Eo, h->P, h->P.center_prev, 4.0, h->ntype[nb],
                                                               nullptr, nullptr, nullptr);
        k_masks_clear<<<pgrid(h, PSLICES), PB, 0, h->stream>>>(h->L, Eo, h->P, h->P.center_old, 4.0,
                                                               h->ntype[nb] /*already clear*/, h->walls_i, h->pin_i,
                                                               h->pout_i);
        build_masks(h, nb);
    }
    static void st_refill(cgl_handle *h)
    {
        const int nb = h->ntcur ^ 1;
        P *phi = (P *)h->phi[h->phicur ^ 1];
        const Ext En = ext_with(h, h->ntype[nb]);
        k_refill<P><<<pgrid(h, LSLICES), PB, 0, h->stream>>>(h->L, En, h->P, h->ntype[h->ntcur], h->ntype[nb],
                                                             h->pout_i, phi);
        mirror_phi(h, phi);
        mirror_mac(h);
    }
    // ghost phi, wetting, derivatives, collision, Recolor, boundary and
    // streaming (one fused kernel), then curve; the new masks take force
    static void st_collide_stream(cgl_handle *h)
    {
        const int nb = h->npar ? h->ntcur ^ 1 : h->ntcur;
        P *phi = (P *)h->phi[h->phicur ^ 1];
        const Ext En = ext_with(h, h->ntype[nb]);
        if (h->npar) {
            k_ghost_phi<P><<<pgrid(h, LSLICES), PB, 0, h->stream>>>(h->L, En, h->P, h->ntype[nb], h->pin_i, phi);
            mirror_phi(h, phi);
            // wetted normals next to the particles, read by the collide kernel
            k_wet_normals<R, P><<<pgrid(h, LSLICES), PB, 0, h->stream>>>(h->L, En, h->P, phi);
        }
        const void *i0, *i1;
        void *o0, *o1;
        state_in(h, i0, i1);
        state_out(h, o0, o1);
        const P *phio = (const P *)h->phi[h->phicur];
        if (h->tev) (void)hipEventRecord(h->tev[2], h->stream);
        const dim3 g2 = k2_grid<R>(h->L, h->xc);
        if (h->state_layout == RB && h->layout == RB)
            k_collide<RB, RB, S, R, P, SHIFT, true>
                <<<g2, k2_block<R>(), 0, h->stream>>>(h->L, i0, i1, o0, o1, phi, phio, h->xc, En);
        else if (h->state_layout == RB)
            k_collide<RB, FS, S, R, P, SHIFT, true>
                <<<g2, k2_block<R>(), 0, h->stream>>>(h->L, i0, i1, o0, o1, phi, phio, h->xc, En);
        else {
            // the tile kernel for all but the special lanes, then those
            k_collide<FS, FS, S, R, P, SHIFT, false, true>
                <<<g2, k2_block<R>(), 0, h->stream>>>(h->L, i0, i1, o0, o1, phi, phio, h->xc, En);
            if (mirrors(h))
                k_collide_special<S, R, P, SHIFT, false>
                    <<<face_grid(h), 256, 0, h->stream>>>(h->L, i0, i1, o0, o1, phi, phio, En, h->P);
            if (h->npar)
                k_collide_special<S, R, P, SHIFT, true><<<dim3(h->band_blocks, h->npar), PB, 0, h->stream>>>(
                    h->L, i0, i1, o0, o1, phi, phio, En, h->P);
        }
        if (h->tev) (void)hipEventRecord(h->tev[3], h->stream);
        h->state_layout = h->layout;
        if (h->layout == FS && ghost_faces(h))
            k_fill_ghosts<S, SHIFT><<<face_grid(h), 256, 0, h->stream>>>(h->L, h->E, (S *)o0, En.rbo[0], En.rbo[1]);
        if (h->npar) {
            if (h->layout == FS)
                k_curve_fs<S, SHIFT><<<pgrid(h, LSLICES), PB, 0, h->stream>>>(h->L, En, h->P, (S *)o0, o1,
                                                                              En.rbo[0], En.rbo[1]);
            else
                k_curve<S, SHIFT><<<pgrid(h, LSLICES), PB, 0, h->stream>>>(h->L, En, h->P, h->ntype[nb], h->walls_i,
                                                                          h->pout_i, (S *)o0, (S *)o1,
                                                                          SHIFT ? (const float *)phi : nullptr);
            if (mirrors(h)) k_mirror_der<<<face_grid(h), 256, 0, h->stream>>>(h->L, h->E, h->E.der);
            h->ntcur = nb;
        }
    }
    // ---- scatter-form step with wall faces and particles (FP32/FP16
    // factorised storage, no mirror faces).  The tile step kernel and the
    // band kernel add the next phase field's fixed-point terms (the band's by
    // atomic addition), the links add theirs (k_curve_fs), and the finishing
    // pass forms the field; the wall, particle and band kernels read and
    // write that buffer as phi through QPhi.  The stages are those of
    // step_ext, with calmacros of the band from the scatter step's field.
    static void step_sc_ext(cgl_handle *h, bool last)
    {
        if constexpr (SHIFT && std::is_same_v<P, float>) {
            QPhi *phi = (QPhi *)h->phi[h->phicur ^ 1];
            if (!h->phi_ok) {
                // the current state's phase field by the gather
                moments_ext_t<QPhi>(h, false);
            } else {
                if (h->npar)
                    k_band_sc<S><<<dim3(h->band_blocks, h->npar), PB, 0, h->stream>>>(
                        h->L, ext_with(h, h->ntype[h->ntcur]), h->P, (const S *)h->fsf[h->cur], (float *)phi);
                wall_phi(h, phi);
            }
            int nb = h->ntcur;
            if (h->npar) {
                const Ext Eo = ext_with(h, h->ntype[h->ntcur]);
                k_lagrange<QPhi><<<pgrid(h, LAG_BLOCKS), PB, 0, h->stream>>>(h->L, Eo, h->P, phi, h->walls_i);
                st_momentum(h);
                st_move(h, MV_ALL);
                st_set_walls(h);
                nb = h->ntcur ^ 1;
                const Ext En = ext_with(h, h->ntype[nb]);
                k_refill<QPhi><<<pgrid(h, LSLICES), PB, 0, h->stream>>>(h->L, En, h->P, h->ntype[h->ntcur],
                                                                       h->ntype[nb], h->pout_i, phi);
                k_ghost_phi<QPhi><<<pgrid(h, LSLICES), PB, 0, h->stream>>>(h->L, En, h->P, h->ntype[nb], h->pin_i, phi);
                k_wet_normals<R, QPhi><<<pgrid(h, LSLICES), PB, 0, h->stream>>>(h->L, En, h->P, phi);
            }
            const Ext En = ext_with(h, h->ntype[nb]);
            const S *fin = (const S *)h->fsf[h->cur];
            S *fout = (S *)h->fsf[h->cur ^ 1];
            float4 *sall = (float4 *)h->fss[h->cur ^ 1];
            Lat Ls = h->L;
            Ls.xcd_remap = h->sc_remap;
            if (h->npar)
                k_band_normals<<<dim3(h->nrm_blocks, h->npar), PB, 0, h->stream>>>(h->L, En, h->P,
                                                                                   (const float *)phi, h->bnrm);
            if (h->tev) (void)hipEventRecord(h->tev[2], h->stream);
            const dim3 grid((h->L.nz + TZ - 1) / TZ, (h->L.ny + KTY<float> - 1) / KTY<float>,
                            (h->L.nx + h->xc - 1) / h->xc);
            // (per-direction bases at every size: the one-base form spills registers here)
            k_collide_sc_ext<S, false><<<grid, dim3(TZ, KTY<float>, 1), 0, h->stream>>>(
                Ls, fin, fout, last ? sall : nullptr, (const float *)phi, h->sb, h->xc, En);
            if (h->npar)
                k_collide_sc_band<S><<<dim3(h->band_blocks, h->npar), PB, 0, h->stream>>>(
                    h->L, En, h->P, fin, fout, sall, (const float *)phi, h->bnrm, h->sb.own);
            if (h->tev) (void)hipEventRecord(h->tev[3], h->stream);
            h->state_layout = FS;
            // (the step kernels have written the wall ghost slots)
            if (h->npar) {
                k_curve_fs<S, SHIFT><<<pgrid(h, LSLICES), PB, 0, h->stream>>>(h->L, En, h->P, fout, sall, En.rbo[0],
                                                                              En.rbo[1], h->sb.own);
                h->ntcur = nb;
            }
            h->scal_stale = !last;
            // timing runs: the finishing pass counts as the step's moments
            // pass, the tile and band kernels as its collide kernels
            if (h->tev) (void)hipEventRecord(h->tev[0], h->stream);
            sc_fin(h);
            if (h->tev) (void)hipEventRecord(h->tev[1], h->stream);
        }
    }
    // one time step with walls and/or particles (S:216-251)
    static void step_ext(cgl_handle *h)
    {
        if (h->npar && !h->primed) {
            prime(h);
            h->primed = 1;
        }
        if (h->tev) (void)hipEventRecord(h->tev[0], h->stream);
        moments_ext(h, false);
        if (h->tev) (void)hipEventRecord(h->tev[1], h->stream);
        if (h->npar) {
            st_lagrange(h);
            st_momentum(h);
            st_move(h, MV_ALL);
            st_set_walls(h);
            st_refill(h);
        }
        st_collide_stream(h);
    }

    static void init_cylinder(cgl_handle *h, const CylIC &ic)
    {
        k_init_phi<P><<<node_grid(h->L), node_block(), 0, h->stream>>>(h->L, ic, (P *)h->phi[h->phicur]);
        k_init_pops<S, P, SHIFT><<<node_grid(h->L), node_block(), 0, h->stream>>>(
            h->L, ic, (S *)h->pops[h->cur][0], (S *)h->pops[h->cur][1], (const P *)h->phi[h->phicur], h->E);
    }
};

using ImplF64 = Impl<double, double, double, false>;
using ImplF32 = Impl<float, float, float, true>;
using ImplF16 = Impl<_Float16, float, float, true>;

} // namespace cgl

extern template struct cgl::Impl<double, double, double, false>;
extern template struct cgl::Impl<float, float, float, true>;
extern template struct cgl::Impl<_Float16, float, float, true>;
