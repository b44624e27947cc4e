// BoundaryManager.hpp - boundary-condition registry with the reference's
// interface (cpp/src/core/BoundaryManager.hpp:8-72). The add_* calls write
// the cell flags of the host Grid at scene set-up; the apply_* calls are the
// host versions for callers that edit a Grid directly. During a FluidSolver
// step the same masks are applied on the GPU, fused into the stencil kernels.
#pragma once
#include <array>
#include <vector>

#include "Grid.hpp"

namespace ffroom {

struct FanBC {
    int i_min = 0, i_max = 0;
    int j_min = 0, j_max = 0;
    int k_min = 0, k_max = 0;
    std::array<double, 3> vel{{0.0, 0.0, 0.0}};
};

struct OutflowBC {
    int axis = 0;  // 0=x, 1=y, 2=z
    int side = 0;  // 0=low face, 1=high face
    int a_min = 0, a_max = 0;
    int b_min = 0, b_max = 0;
};

struct OpeningBC : OutflowBC {
    double T_outside = 20.0;
};

class BoundaryManager {
public:
    void add_solid_box(Grid &grid, int i0, int i1, int j0, int j1, int k0, int k1);
    void add_fan(Grid &grid, const FanBC &fan);
    void add_outflow(Grid &grid, const OutflowBC &obc);
    void add_opening(Grid &grid, const OpeningBC &obc);

    void apply_noslip(Grid &grid);
    void apply_inflow(Grid &grid);
    void apply_outflow(Grid &grid);
    void apply_opening_temperature(Grid &grid);

    const std::vector<FanBC> &fans() const { return fans_; }
    const std::vector<OutflowBC> &outflows() const { return outflows_; }
    const std::vector<OpeningBC> &openings() const { return openings_; }

    // Optional inlet temperature of fan `index` (no counterpart in the
    // reference, whose INFLOW cells keep T_initial): FluidSolver sets the
    // fan's cells to this temperature at construction.
    void set_fan_temperature(int index, double T_inlet);
    const std::vector<std::pair<int, double>> &fan_temperatures() const { return fan_T_; }

    uint64_t version() const { return version_; }

private:
    std::vector<FanBC> fans_;
    std::vector<OutflowBC> outflows_;
    std::vector<OpeningBC> openings_;
    std::vector<std::pair<int, double>> fan_T_;
    uint64_t version_ = 0;
};

}  // namespace ffroom
