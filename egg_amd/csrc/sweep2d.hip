// 2D instantiations of the structured sweep kernel (sweep_impl.hpp): quad
// patches of 4 cells x 4 corner samples, shape / untangle / shape_size
// metrics; tile shapes in nodes (last axis x first axis). Variants 0-3: one
// lane per cell (4 per node); 4-5: one lane per sample (16 per node); 6-7:
// one lane per cell without LDS staging; 8, 11: two cells per lane (2 lanes
// per node), 10: the same without LDS staging; 9: one lane per node (its 4
// cells in turn); 12, 13: variant 0 with 4 / 8 lanes per stored DOF instead
// of 16 (stored occurrences in turn). The one-lane-per-cell
// shapes give the same bits; the cells-per-lane ones fold the cells in the
// same order but the compiler fuses their multiply-adds differently.
#include "sweep_impl.hpp"

#include <stdexcept>
#include <string>

namespace eggmi
{

template <int KIND, bool TAU, bool GRAD, bool PROP = false> static void sweep2d_v(int v, const SweepArgs& a, hipStream_t s)
{
    if constexpr (!PROP) {  // no proposal kernels for these
        if (v == 4) return launch_sample_variant<KIND, 16, 1, TAU, GRAD>(a, s);  // one lane per sample
        if (v == 5) return launch_sample_variant<KIND, 8, 2, TAU, GRAD>(a, s);
        if (v == 6) return launch_variant<2, KIND, 64, 1, 1, TAU, GRAD, false, false>(a, s);  // no LDS staging
        if (v == 7) return launch_variant<2, KIND, 32, 2, 1, TAU, GRAD, false, false>(a, s);
        if (v == 8) return launch_variant<2, KIND, 32, 4, 1, TAU, GRAD, false, true, 2>(a, s);  // cells per lane
        if (v == 9) return launch_variant<2, KIND, 64, 4, 1, TAU, GRAD, false, true, 4>(a, s);
        if (v == 10) return launch_variant<2, KIND, 32, 4, 1, TAU, GRAD, false, false, 2>(a, s);  // + no LDS
        if (v == 11) return launch_variant<2, KIND, 64, 2, 1, TAU, GRAD, false, true, 2>(a, s);
        // variant 0's tiles with 4 / 8 lanes per stored DOF (4 / 2 occurrences per lane)
        if (v == 12) return launch_variant<2, KIND, 32, 2, 1, TAU, GRAD, false, true, 1, 4>(a, s);
        if (v == 13) return launch_variant<2, KIND, 32, 2, 1, TAU, GRAD, false, true, 1, 8>(a, s);
    }
    switch (v) {
    case 0: launch_variant<2, KIND, 32, 2, 1, TAU, GRAD, PROP>(a, s); break;
    case 1: launch_variant<2, KIND, 16, 4, 1, TAU, GRAD, PROP>(a, s); break;
    case 2: launch_variant<2, KIND, 64, 1, 1, TAU, GRAD, PROP>(a, s); break;
    case 3: launch_variant<2, KIND, 32, 4, 1, TAU, GRAD, PROP>(a, s); break;
    default: throw std::invalid_argument("2D tile variant " + std::to_string(v) + " has no proposal kernel");
    }
}

// The special stored DOFs (constrained, or carrying curvature windows), laid
// out for the variant's stored lane groups.
template <int KIND, bool GRAD = false> static void launch_constrained(int v, const SweepArgs& a, hipStream_t s)
{
    switch (tile_variant_stored_lanes(2, v)) {
    case 4: constrained_g<2, KIND, 4, GRAD>(a, s); break;
    case 8: constrained_g<2, KIND, 8, GRAD>(a, s); break;
    default: constrained_g<2, KIND, 16, GRAD>(a, s); break;
    }
}

static void sweep2d_main(int kind, int v, const SweepArgs& a, hipStream_t s);

// One workgroup smoothing a whole FAS coarse level (tau-shifted sweeps; the
// tile variant's shape and lane layout)
template <int KIND>
static void level2d_v(int v, const SweepArgs& a, float* x0, float* x1, int n, hipStream_t s)
{
    switch (v) {
    case 0: return launch_level_variant<2, KIND, 32, 2, 1>(a, x0, x1, n, s);
    case 1: return launch_level_variant<2, KIND, 16, 4, 1>(a, x0, x1, n, s);
    case 2: return launch_level_variant<2, KIND, 64, 1, 1>(a, x0, x1, n, s);
    case 3: return launch_level_variant<2, KIND, 32, 4, 1>(a, x0, x1, n, s);
    case 4: return launch_level_variant<2, KIND, 16, 1, 1, true, 1, true>(a, x0, x1, n, s);
    case 5: return launch_level_variant<2, KIND, 8, 2, 1, true, 1, true>(a, x0, x1, n, s);
    case 6: return launch_level_variant<2, KIND, 64, 1, 1, false>(a, x0, x1, n, s);
    case 7: return launch_level_variant<2, KIND, 32, 2, 1, false>(a, x0, x1, n, s);
    case 8: return launch_level_variant<2, KIND, 32, 4, 1, true, 2>(a, x0, x1, n, s);
    case 9: return launch_level_variant<2, KIND, 64, 4, 1, true, 4>(a, x0, x1, n, s);
    case 10: return launch_level_variant<2, KIND, 32, 4, 1, false, 2>(a, x0, x1, n, s);
    case 11: return launch_level_variant<2, KIND, 64, 2, 1, true, 2>(a, x0, x1, n, s);
    case 12:  // variants 12 and 13 differ from 0 only in their stored lanes, which coarse levels lack
    case 13: return launch_level_variant<2, KIND, 32, 2, 1>(a, x0, x1, n, s);
    default: throw std::invalid_argument("2D tile variant " + std::to_string(v) + " out of range");
    }
}

void launch_level_smooth2d(int kind, int v, const SweepArgs& a, float* x0, float* x1, int n, hipStream_t s)
{
    if (kind == kShapeSize)
        level2d_v<kShapeSize>(v, a, x0, x1, n, s);
    else
        level2d_v<kShape>(v, a, x0, x1, n, s);
}

void launch_sweep2d(int kind, int v, const SweepArgs& a, hipStream_t s)
{
    sweep2d_main(kind, v, a, s);
    if (a.con_n > 0 && !a.tau) {
        if (kind == kUntangle)
            launch_constrained<kUntangle>(v, a, s);
        else if (kind == kShapeSize)
            launch_constrained<kShapeSize>(v, a, s);
        else
            launch_constrained<kShape>(v, a, s);
    }
}

static void sweep2d_main(int kind, int v, const SweepArgs& a, hipStream_t s)
{
    if (a.tau) {
        if (kind == kShapeSize)
            sweep2d_v<kShapeSize, true, false>(v, a, s);
        else
            sweep2d_v<kShape, true, false>(v, a, s);  // FAS levels: barrier phases only
        return;
    }
    if (a.mlp.fwd && kind != kUntangle) {  // learned proposal (barrier phases, no FAS shift)
        if (kind == kShapeSize)
            sweep2d_v<kShapeSize, false, false, true>(v, a, s);
        else
            sweep2d_v<kShape, false, false, true>(v, a, s);
        return;
    }
    if (kind == kUntangle)
        sweep2d_v<kUntangle, false, false>(v, a, s);
    else if (kind == kShapeSize)
        sweep2d_v<kShapeSize, false, false>(v, a, s);
    else
        sweep2d_v<kShape, false, false>(v, a, s);
}

static void grad2d_main(int kind, int v, const SweepArgs& a, hipStream_t s);

void launch_grad2d(int kind, int v, const SweepArgs& a, hipStream_t s)
{
    grad2d_main(kind, v, a, s);
    if (a.con_n > 0 && !a.tau) {  // the fine level's special DOFs (FAS gradient)
        if (kind == kShapeSize)
            launch_constrained<kShapeSize, true>(v, a, s);
        else
            launch_constrained<kShape, true>(v, a, s);
    }
}

static void grad2d_main(int kind, int v, const SweepArgs& a, hipStream_t s)
{
    if (a.tau) {
        if (kind == kShapeSize)
            sweep2d_v<kShapeSize, true, true>(v, a, s);
        else
            sweep2d_v<kShape, true, true>(v, a, s);
    } else {
        if (kind == kShapeSize)
            sweep2d_v<kShapeSize, false, true>(v, a, s);
        else
            sweep2d_v<kShape, false, true>(v, a, s);
    }
}

}  // namespace eggmi
