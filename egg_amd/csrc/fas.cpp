// FAS nonlinear multigrid driver (placeholder until the hierarchy lands).
#include "session.hpp"

namespace eggmi
{

std::vector<std::vector<std::array<int, 3>>> Session::mg_level_shapes(int) { return {}; }

std::pair<std::vector<double>, std::vector<double>> Session::run_fas(int, int, int, int, double, int, int)
{
    throw Internal("run_fas: not available in this build");
}

}  // namespace eggmi
