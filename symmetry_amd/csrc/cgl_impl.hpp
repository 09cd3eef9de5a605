// cgl_impl.hpp -- device handle and per-precision kernel launchers.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/cgl.h"
#include "cgl_kernels.hpp"

struct cgl_handle {
    cgl_params p;
    cgl::Lat L;
    size_t N;
    size_t esize;     // bytes per stored population value
    size_t phisize;   // bytes per phi value
    void *pops[2][2]; // [buffer][colour]: colour 0 red, 1 blue; 19*N values each
    void *phi[2];
    int cur;          // pops buffer holding the current state
    int phicur;       // phi buffer holding the shift reference of the current state
    int xc;           // x planes per workgroup of the fused kernel
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

template <class S, class R, class P, bool SHIFT> struct Impl {
    static void moments(cgl_handle *h, bool full)
    {
        const S *rs = (const S *)h->pops[h->cur][0], *bs = (const S *)h->pops[h->cur][1];
        const P *phio = (const P *)h->phi[h->phicur];
        P *phi = (P *)h->phi[h->phicur ^ 1];
        if (full)
            k_moments<S, R, P, SHIFT, true>
                <<<node_grid(h->L), node_block(), 0, h->stream>>>(h->L, rs, bs, phio, phi, h->mac);
        else
            k_moments<S, R, P, SHIFT, false>
                <<<node_grid(h->L), node_block(), 0, h->stream>>>(h->L, rs, bs, phio, phi, nullptr);
    }
    static void collide(cgl_handle *h)
    {
        const S *rin = (const S *)h->pops[h->cur][0], *bin = (const S *)h->pops[h->cur][1];
        S *rout = (S *)h->pops[h->cur ^ 1][0], *bout = (S *)h->pops[h->cur ^ 1][1];
        const P *phio = (const P *)h->phi[h->phicur];
        const P *phi = (const P *)h->phi[h->phicur ^ 1];
        dim3 grid((h->L.nz + TZ - 1) / TZ, (h->L.ny + TY - 1) / TY, (h->L.nx + h->xc - 1) / h->xc);
        k_collide<S, R, P, SHIFT><<<grid, node_block(), 0, h->stream>>>(h->L, rin, bin, rout, bout, phi, phio, h->xc);
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
        k_unpack<S, P, SHIFT><<<node_grid(h->L), node_block(), 0, h->stream>>>(
            h->L, (const S *)h->pops[h->cur][red ? 0 : 1], aos, (const P *)h->phi[h->phicur], red);
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
