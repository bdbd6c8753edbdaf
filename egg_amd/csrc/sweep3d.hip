// 3D instantiations of the structured sweep kernel (sweep_impl.hpp): hex
// patches of 8 cells x 8 corner samples, mu_cond3 family metrics; tile
// shapes in nodes (last axis x middle x first axis). Variants 0-1: one lane
// per cell (8 per node); 2: the same without LDS staging; 3: two cells per
// lane (4 lanes per node): same fold order as one per lane, not the same
// multiply-add fusion; 4: variant 2 with 16 lanes per stored DOF instead of
// 64 (4 occurrences per lane in turn).
#include "sweep_impl.hpp"

#include <stdexcept>
#include <string>

namespace eggmi
{

template <int KIND, bool TAU, bool GRAD, bool PROP = false> static void sweep3d_v(int v, const SweepArgs& a, hipStream_t s)
{
    if constexpr (!PROP) {
        if (v == 2) return launch_variant<3, KIND, 8, 4, 2, TAU, GRAD, false, false>(a, s);  // no LDS staging
        if (v == 3) return launch_variant<3, KIND, 8, 4, 2, TAU, GRAD, false, true, 2>(a, s);  // cells per lane
        if (v == 4)  // no LDS staging, 16 lanes per stored DOF (4 occurrences per lane)
            return launch_variant<3, KIND, 8, 4, 2, TAU, GRAD, false, false, 1, 16>(a, s);
    }
    switch (v) {
    case 0: launch_variant<3, KIND, 8, 4, 2, TAU, GRAD, PROP>(a, s); break;
    case 1: launch_variant<3, KIND, 4, 4, 4, TAU, GRAD, PROP>(a, s); break;
    default: throw std::invalid_argument("3D tile variant " + std::to_string(v) + " has no proposal kernel");
    }
}

// The special stored DOFs (those with directional windows), laid out for the
// variant's stored lane groups.
template <int KIND, bool GRAD = false> static void launch_special3d(int v, const SweepArgs& a, hipStream_t s)
{
    if (tile_variant_stored_lanes(3, v) == 16)
        constrained_g<3, KIND, 16, GRAD>(a, s);
    else
        constrained_g<3, KIND, 64, GRAD>(a, s);
}

static void sweep3d_main(int kind, int v, const SweepArgs& a, hipStream_t s);

// One workgroup smoothing a whole FAS coarse level (tau-shifted sweeps)
template <int KIND>
static void level3d_v(int v, const SweepArgs& a, float* x0, float* x1, int n, hipStream_t s)
{
    switch (v) {
    case 0: return launch_level_variant<3, KIND, 8, 4, 2>(a, x0, x1, n, s);
    case 1: return launch_level_variant<3, KIND, 4, 4, 4>(a, x0, x1, n, s);
    case 2:  // variant 4 differs from 2 only in its stored lanes, which coarse levels lack
    case 4: return launch_level_variant<3, KIND, 8, 4, 2, false>(a, x0, x1, n, s);
    case 3: return launch_level_variant<3, KIND, 8, 4, 2, true, 2>(a, x0, x1, n, s);
    default: throw std::invalid_argument("3D tile variant " + std::to_string(v) + " out of range");
    }
}

void launch_level_smooth3d(int kind, int v, const SweepArgs& a, float* x0, float* x1, int n, hipStream_t s)
{
    if (kind == kShapeSize)
        level3d_v<kShapeSize>(v, a, x0, x1, n, s);
    else
        level3d_v<kShape>(v, a, x0, x1, n, s);
}
static void grad3d_main(int kind, int v, const SweepArgs& a, hipStream_t s);

void launch_sweep3d(int kind, int v, const SweepArgs& a, hipStream_t s)
{
    sweep3d_main(kind, v, a, s);
    if (a.con_n > 0 && !a.tau) {
        if (kind == kUntangle)
            launch_special3d<kUntangle>(v, a, s);
        else if (kind == kShapeSize)
            launch_special3d<kShapeSize>(v, a, s);
        else
            launch_special3d<kShape>(v, a, s);
    }
}

void launch_grad3d(int kind, int v, const SweepArgs& a, hipStream_t s)
{
    grad3d_main(kind, v, a, s);
    if (a.con_n > 0 && !a.tau) {  // the fine level's special DOFs (FAS gradient)
        if (kind == kShapeSize)
            launch_special3d<kShapeSize, true>(v, a, s);
        else
            launch_special3d<kShape, true>(v, a, s);
    }
}

static void sweep3d_main(int kind, int v, const SweepArgs& a, hipStream_t s)
{
    if (a.tau) {
        if (kind == kShapeSize)
            sweep3d_v<kShapeSize, true, false>(v, a, s);
        else
            sweep3d_v<kShape, true, false>(v, a, s);  // FAS levels: barrier phases only
        return;
    }
    if (a.mlp.fwd && kind != kUntangle) {  // learned proposal (barrier phases, no FAS shift)
        if (kind == kShapeSize)
            sweep3d_v<kShapeSize, false, false, true>(v, a, s);
        else
            sweep3d_v<kShape, false, false, true>(v, a, s);
        return;
    }
    if (kind == kUntangle)
        sweep3d_v<kUntangle, false, false>(v, a, s);
    else if (kind == kShapeSize)
        sweep3d_v<kShapeSize, false, false>(v, a, s);
    else
        sweep3d_v<kShape, false, false>(v, a, s);
}

static void grad3d_main(int kind, int v, const SweepArgs& a, hipStream_t s)
{
    if (a.tau) {
        if (kind == kShapeSize)
            sweep3d_v<kShapeSize, true, true>(v, a, s);
        else
            sweep3d_v<kShape, true, true>(v, a, s);
    } else {
        if (kind == kShapeSize)
            sweep3d_v<kShapeSize, false, true>(v, a, s);
        else
            sweep3d_v<kShape, false, true>(v, a, s);
    }
}

}  // namespace eggmi
