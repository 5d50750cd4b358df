// sites.hpp -- device helpers shared by the Hamiltonian build (k_sites.hip)
// and the map-coefficient gradient (adjoint.hip).
#pragma once

#include "common.hpp"

namespace cir {

template <typename T>
struct MapT {
    T w0, w1, w2, x0, x1, p0, p1, m0, m1, m2, wi0, wi1, wi2;
};

template <typename T>
MapT<T> map_t(const MapCoef& c) {
    return MapT<T>{T(c.w[0]), T(c.w[1]), T(c.w[2]), T(c.x[0]), T(c.x[1]), T(c.p[0]), T(c.p[1]),
                   T(c.mup[0]), T(c.mup[1]), T(c.mup[2]), T(c.wi[0]), T(c.wi[1]), T(c.wi[2])};
}

// Transition-dipole coupling geometry of chromophores i and j
// (calcIR.cu:1353-1401): (u_i.u_j - 3 (u_i.d)(u_j.d)) / (r in bohr)^3 with
// d the unit min-image separation of the TDM positions.  The coupling is
// this times (x_i mu'_i)(x_j mu'_j) times the hartree -> cm^-1 factor.  The
// operation order is symmetric in i and j, so H(i,j) == H(j,i) bit for bit.
template <typename T>
__device__ __forceinline__ T tdc_geometry(T uix, T uiy, T uiz, T rix, T riy, T riz, T ujx, T ujy, T ujz, T rjx,
                                          T rjy, T rjz, const PBox<T>& bb) {
    T dr[3] = {rix - rjx, riy - rjy, riz - rjz};
    minimage3<T>(dr[0], dr[1], dr[2], bb);
    const T r = sqrt(dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]);
    const T inv = T(1) / r;
    const T d0 = dr[0] * inv, d1 = dr[1] * inv, d2 = dr[2] * inv;
    const T uu = uix * ujx + uiy * ujy + uiz * ujz;
    const T ur = uix * d0 + uiy * d1 + uiz * d2;
    const T vr = ujx * d0 + ujy * d1 + ujz * d2;
    const T rb = r * T(kBohrPerNm);
    const T t = ur * vr;
    return (uu - T(3) * t) / (rb * rb * rb);
}

}  // namespace cir
