// cgl_lattice.hpp -- D3Q19 lattice constants and the MRT moment basis as
// compile-time tables, so that fully unrolled device loops fold every zero
// coefficient away and every +-1 into an add or subtract.
//
// Direction order, weights and the moment basis are the reference's
// (Source1.f90 S:105-142); they fix the host array layout (0:18, zl, yl, xl)
// that the C ABI accepts, so they have to be the same tables.
#pragma once

#include <type_traits>

namespace cgl {

constexpr int Q = 19;
constexpr int CX[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
constexpr int CY[Q] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
constexpr int CZ[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
constexpr int OPC[Q] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};
constexpr double W[Q] = {1.0 / 3.0,  1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0,
                         1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
                         1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};

// Moment k of direction i for the d'Humieres D3Q19 basis:
// rho, e, eps, jx, qx, jy, qy, jz, qz, 3pxx, 3pixx, pww, piww, pxy, pyz, pxz, mx, my, mz.
constexpr int moment(int k, int i)
{
    int cx = CX[i], cy = CY[i], cz = CZ[i];
    int c2 = cx * cx + cy * cy + cz * cz;
    switch (k) {
    case 0: return 1;
    case 1: return 19 * c2 - 30;
    case 2: return (21 * c2 * c2 - 53 * c2 + 24) / 2;
    case 3: return cx;
    case 4: return (5 * c2 - 9) * cx;
    case 5: return cy;
    case 6: return (5 * c2 - 9) * cy;
    case 7: return cz;
    case 8: return (5 * c2 - 9) * cz;
    case 9: return 3 * cx * cx - c2;
    case 10: return (3 * c2 - 5) * (3 * cx * cx - c2);
    case 11: return cy * cy - cz * cz;
    case 12: return (3 * c2 - 5) * (cy * cy - cz * cz);
    case 13: return cx * cy;
    case 14: return cy * cz;
    case 15: return cx * cz;
    case 16: return cx * (cy * cy - cz * cz);
    case 17: return cy * (cz * cz - cx * cx);
    default: return cz * (cx * cx - cy * cy);
    }
}

struct MTab {
    int m[Q][Q];
    double norm2[Q];
    double inv[Q][Q]; // inv[i][k] = m[k][i] / norm2[k]
    bool even[Q];     // moment k takes equal values on opposite directions
};

constexpr MTab make_mtab()
{
    MTab t{};
    for (int k = 0; k < Q; ++k) {
        double n = 0.0;
        for (int i = 0; i < Q; ++i) {
            t.m[k][i] = moment(k, i);
            n += double(t.m[k][i]) * double(t.m[k][i]);
        }
        t.norm2[k] = n;
    }
    for (int i = 0; i < Q; ++i)
        for (int k = 0; k < Q; ++k) t.inv[i][k] = double(t.m[k][i]) / t.norm2[k];
    for (int k = 0; k < Q; ++k) {
        t.even[k] = true;
        for (int i = 0; i < Q; ++i)
            if (t.m[k][OPC[i]] != t.m[k][i]) t.even[k] = false;
    }
    return t;
}

constexpr MTab MT = make_mtab();

// Direction pairs (p, opc(p)) with p the first of each pair.  With the moment
// parities MT.even they halve the work of the MRT transforms.
constexpr int NPAIR = 9;
constexpr int PAIR[NPAIR] = {1, 3, 5, 7, 8, 11, 12, 15, 16};

// REFL[A][i]: direction i with the velocity components of the axes in mask A
// (bit 0 x, bit 1 y, bit 2 z) reversed -- specular reflection in mirror faces.
struct ReflTab {
    int d[8][Q];
};
constexpr int find_dir(int cx, int cy, int cz)
{
    for (int i = 0; i < Q; ++i)
        if (CX[i] == cx && CY[i] == cy && CZ[i] == cz) return i;
    return -1;
}
constexpr ReflTab make_refl()
{
    ReflTab t{};
    for (int A = 0; A < 8; ++A)
        for (int i = 0; i < Q; ++i)
            t.d[A][i] = find_dir((A & 1) ? -CX[i] : CX[i], (A & 2) ? -CY[i] : CY[i], (A & 4) ? -CZ[i] : CZ[i]);
    return t;
}
constexpr ReflTab REFL = make_refl();

// Compile-time loop: f(std::integral_constant<int, 0>) ... f(<N-1>).
template <int N, class F, int K = 0> __host__ __device__ constexpr void static_for(F &&f)
{
    if constexpr (K < N) {
        f(std::integral_constant<int, K>{});
        static_for<N, F, K + 1>(static_cast<F &&>(f));
    }
}

// Moments of the rest-state weights: M w = (1, -11, 3, 0, ..., 0).
constexpr double MW0 = 1.0, MW1 = -11.0, MW2 = 3.0;

} // namespace cgl
