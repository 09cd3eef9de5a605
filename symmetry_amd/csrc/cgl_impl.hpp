// cgl_impl.hpp -- device handle and per-precision kernel launchers.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "../../include/cgl.h"
#include "cgl_kernels.hpp"

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
    int tune_best_xc;
    float tune_best_ms;
    double *mac;      // scratch: 7*N doubles for macroscopic downloads (lazy)
    double *aos;      // scratch: 19*N doubles for host-layout conversion (lazy)
    int *iflag;       // scratch: N ints (lazy)
    hipStream_t stream;
    hipEvent_t ev[3];
    int steps_done;
};

namespace cgl {

inline dim3 node_grid(const Lat &L) { return dim3((L.nz + TZ - 1) / TZ, (L.ny + TY - 1) / TY, L.nx); }
inline dim3 node_block() { return dim3(TZ, TY, 1); }

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

template <class S, class R, class P, bool SHIFT> struct Impl {
    static void moments(cgl_handle *h, bool full)
    {
        const void *i0, *i1;
        state_in(h, i0, i1);
        const P *phio = (const P *)h->phi[h->phicur];
        P *phi = (P *)h->phi[h->phicur ^ 1];
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
        if (h->state_layout == RB)
            launch(std::integral_constant<int, RB>{});
        else
            launch(std::integral_constant<int, FS>{});
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
        dim3 grid((h->L.nz + TZ - 1) / TZ, (h->L.ny + TY - 1) / TY, (h->L.nx + h->xc - 1) / h->xc);
        if (h->state_layout == RB && h->layout == RB)
            k_collide<RB, RB, S, R, P, SHIFT>
                <<<grid, node_block(), 0, h->stream>>>(h->L, i0, i1, o0, o1, phi, phio, h->xc);
        else if (h->state_layout == RB)
            k_collide<RB, FS, S, R, P, SHIFT>
                <<<grid, node_block(), 0, h->stream>>>(h->L, i0, i1, o0, o1, phi, phio, h->xc);
        else
            k_collide<FS, FS, S, R, P, SHIFT>
                <<<grid, node_block(), 0, h->stream>>>(h->L, i0, i1, o0, o1, phi, phio, h->xc);
        h->state_layout = h->layout;
    }
    // derivatives of the phi written by the last moments() call
    static void derivatives(cgl_handle *h, double *out, int *flag)
    {
        k_derivatives<P><<<node_grid(h->L), node_block(), 0, h->stream>>>(h->L, (const P *)h->phi[h->phicur ^ 1], out,
                                                                         flag);
    }
    static void phi_aos(cgl_handle *h, const double *ra, const double *ba)
    {
        unsigned nb = (unsigned)((h->N + 255) / 256);
        k_phi_aos<P><<<nb, 256, 0, h->stream>>>(h->L, ra, ba, (P *)h->phi[h->phicur]);
    }
    static void pack(cgl_handle *h, const double *aos, int red)
    {
        k_pack<S, P, SHIFT><<<node_grid(h->L), node_block(), 0, h->stream>>>(
            h->L, aos, (S *)h->pops[h->cur][red ? 0 : 1], (const P *)h->phi[h->phicur], red);
    }
    static void unpack(cgl_handle *h, double *aos, int red)
    {
        if (h->state_layout == RB)
            k_unpack<S, P, SHIFT><<<node_grid(h->L), node_block(), 0, h->stream>>>(
                h->L, (const S *)h->pops[h->cur][red ? 0 : 1], aos, (const P *)h->phi[h->phicur], red);
        else
            k_unpack_fs<S, SHIFT><<<node_grid(h->L), node_block(), 0, h->stream>>>(
                h->L, (const S *)h->fsf[h->cur], h->fss[h->cur], aos, red);
    }
    static void init_cylinder(cgl_handle *h, const CylIC &ic)
    {
        k_init_phi<P><<<node_grid(h->L), node_block(), 0, h->stream>>>(h->L, ic, (P *)h->phi[h->phicur]);
        k_init_pops<S, P, SHIFT><<<node_grid(h->L), node_block(), 0, h->stream>>>(
            h->L, ic, (S *)h->pops[h->cur][0], (S *)h->pops[h->cur][1], (const P *)h->phi[h->phicur]);
    }
};

using ImplF64 = Impl<double, double, double, false>;
using ImplF32 = Impl<float, float, float, true>;
using ImplF16 = Impl<_Float16, float, float, true>;

} // namespace cgl

extern template struct cgl::Impl<double, double, double, false>;
extern template struct cgl::Impl<float, float, float, true>;
extern template struct cgl::Impl<_Float16, float, float, true>;
