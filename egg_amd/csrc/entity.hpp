// 2D closed-form entities a constrained boundary DOF slides on (reference
// src/geometry.hpp:288-320 LineSeg, :322-357 LineRail, :398-431 Circle,
// :432-464 Ellipse): tag 1 segment (sx, sy, ex, ey), 2 circle (cx, cy, r),
// 3 ellipse (cx, cy, rx, ry; radial-scaling projection), 15 rail (a segment
// that releases the node beyond its ends). Shared by the sweep kernels (R =
// float, the FP32 tolerances of src/core.hpp: znorm 1e-7, tiny 1e-20) and
// the host geometry calls (R = double: 1e-15, 1e-30).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace eggmi
{

template <class R> struct EntityTol;
template <> struct EntityTol<float> {
    static constexpr float znorm = 1e-7f, tiny = 1e-20f, inv_sqrt2 = 0.70710678f;
};
template <> struct EntityTol<double> {
    static constexpr double znorm = 1e-15, tiny = 1e-30, inv_sqrt2 = 0.70710678118654752440;
};

template <class R> __host__ __device__ inline R ent_sqrt(R v) { return std::sqrt(v); }
template <class R> __host__ __device__ inline R ent_clamp01(R v) { return v < R(0) ? R(0) : (v > R(1) ? R(1) : v); }

// foot parameter of p on the line through the segment (0 at its start, 1 at its end)
template <class R> __host__ __device__ inline R rail_foot(const R* e, const R* p)
{
    const R abx = e[2] - e[0], aby = e[3] - e[1];
    const R ab2 = abx * abx + aby * aby;
    return ((p[0] - e[0]) * abx + (p[1] - e[1]) * aby) / (ab2 > EntityTol<R>::tiny ? ab2 : EntityTol<R>::tiny);
}

template <class R> __host__ __device__ inline void ent_project(int tag, const R* e, const R* p, R* out)
{
    using T = EntityTol<R>;
    if (tag == 1 || tag == 15) {
        const R abx = e[2] - e[0], aby = e[3] - e[1];
        R t = rail_foot(e, p);
        const bool released = tag == 15 && (t < R(0) || t > R(1));
        t = ent_clamp01(t);
        out[0] = released ? p[0] : e[0] + t * abx;
        out[1] = released ? p[1] : e[1] + t * aby;
    } else if (tag == 3) {
        const R sx = (p[0] - e[0]) / e[2], sy = (p[1] - e[1]) / e[3];
        const R dist = ent_sqrt(sx * sx + sy * sy);
        out[0] = e[0] + (dist < T::znorm ? T::inv_sqrt2 : sx / dist) * e[2];
        out[1] = e[1] + (dist < T::znorm ? T::inv_sqrt2 : sy / dist) * e[3];
    } else {
        const R dx = p[0] - e[0], dy = p[1] - e[1];
        const R dist = ent_sqrt(dx * dx + dy * dy);
        out[0] = dist < T::znorm ? e[0] + e[2] : e[0] + e[2] * dx / dist;
        out[1] = dist < T::znorm ? e[1] : e[1] + e[2] * dy / dist;
    }
}

template <class R> __host__ __device__ inline void ent_tangent(int tag, const R* e, const R* p, R* t)
{
    using T = EntityTol<R>;
    if (tag == 3) {
        // (-rx sin a, ry cos a) at the parametric angle a of the radially
        // scaled point: sin a, cos a as its coordinates over its length
        const R sx = (p[0] - e[0]) / e[2], sy = (p[1] - e[1]) / e[3];
        const R dist = ent_sqrt(sx * sx + sy * sy);
        const R sa = dist < T::znorm ? R(0) : sy / dist, ca = dist < T::znorm ? R(1) : sx / dist;
        const R tx = -e[2] * sa, ty = e[3] * ca;
        const R n = ent_sqrt(tx * tx + ty * ty);
        t[0] = n < T::znorm ? R(1) : tx / n;
        t[1] = n < T::znorm ? R(0) : ty / n;
        return;
    }
    if (tag == 1 || tag == 15) {
        const R abx = e[2] - e[0], aby = e[3] - e[1];
        const R n = ent_sqrt(abx * abx + aby * aby);
        t[0] = n < T::znorm ? R(1) : abx / n;
        t[1] = n < T::znorm ? R(0) : aby / n;
    } else {
        const R dx = p[0] - e[0], dy = p[1] - e[1];
        const R rn = ent_sqrt(dx * dx + dy * dy);
        t[0] = rn < T::znorm ? -R(0) : -dy / rn;
        t[1] = rn < T::znorm ? R(1) : dx / rn;
    }
}

}  // namespace eggmi
