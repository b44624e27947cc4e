// BoundaryManager.cpp - set-up flag writers and host mask helpers
// (reference: cpp/src/core/BoundaryManager.cpp:6-220).
#include "BoundaryManager.hpp"

#include <algorithm>
#include <stdexcept>

namespace ffroom {

namespace {
// Host helpers touch the Grid's vectors: make them current first and mark
// them as the newer copy afterwards.
struct HostEdit {
    Grid &g;
    explicit HostEdit(Grid &g_) : g(g_) { g.sync_to_host(); }
    ~HostEdit() { g.mark_host_modified(); }
};
}  // namespace

void BoundaryManager::add_solid_box(Grid &g, int i0, int i1, int j0, int j1, int k0, int k1)
{   // :6-14, clamped to the grid
    for (int i = std::max(0, i0); i <= std::min(g.Nx - 1, i1); i++)
        for (int j = std::max(0, j0); j <= std::min(g.Ny - 1, j1); j++)
            for (int k = std::max(0, k0); k <= std::min(g.Nz - 1, k1); k++)
                g.cell_type[g.c_idx(i, j, k)] = CellType::SOLID;
    g.geometry_version++;
}

void BoundaryManager::add_fan(Grid &g, const FanBC &f)
{   // :16-22; the fan is registered even if its cells are re-flagged later
    fans_.push_back(f);
    for (int i = std::max(0, f.i_min); i <= std::min(g.Nx - 1, f.i_max); i++)
        for (int j = std::max(0, f.j_min); j <= std::min(g.Ny - 1, f.j_max); j++)
            for (int k = std::max(0, f.k_min); k <= std::min(g.Nz - 1, f.k_max); k++)
                g.cell_type[g.c_idx(i, j, k)] = CellType::INFLOW;
    version_++;
    g.geometry_version++;
}

void BoundaryManager::add_outflow(Grid &g, const OutflowBC &o)
{   // :24-44; unlike the reference the patch must lie on the wall
    const int na = o.axis == 0 ? g.Ny : g.Nx, nb = o.axis == 2 ? g.Ny : g.Nz;
    if (o.axis < 0 || o.axis > 2 || o.side < 0 || o.side > 1 || o.a_min < 0 || o.b_min < 0 ||
        o.a_max >= na || o.b_max >= nb)
        throw std::invalid_argument("outflow patch outside the domain wall");
    outflows_.push_back(o);
    for (int a = o.a_min; a <= o.a_max; a++)
        for (int b = o.b_min; b <= o.b_max; b++) {
            int c;
            if (o.axis == 0) c = g.c_idx(o.side ? g.Nx - 1 : 0, a, b);
            else if (o.axis == 1) c = g.c_idx(a, o.side ? g.Ny - 1 : 0, b);
            else c = g.c_idx(a, b, o.side ? g.Nz - 1 : 0);
            g.cell_type[c] = CellType::OUTFLOW;
        }
    version_++;
    g.geometry_version++;
}

void BoundaryManager::add_opening(Grid &g, const OpeningBC &o)
{   // :132-135
    openings_.push_back(o);
    add_outflow(g, static_cast<const OutflowBC &>(o));
}

void BoundaryManager::set_fan_temperature(int index, double T_inlet)
{
    if (index < 0 || index >= (int)fans_.size()) throw std::out_of_range("fan index");
    fan_T_.emplace_back(index, T_inlet);
    version_++;
}

void BoundaryManager::apply_noslip(Grid &g)
{   // :60-109
    HostEdit he(g);
    auto outflow = [&](int i, int j, int k) { return g.cell_type[g.c_idx(i, j, k)] == CellType::OUTFLOW; };
    for (int j = 0; j < g.Ny; j++)
        for (int k = 0; k < g.Nz; k++) {
            if (!outflow(0, j, k)) g.u[g.u_idx(0, j, k)] = 0.0;
            if (!outflow(g.Nx - 1, j, k)) g.u[g.u_idx(g.Nx, j, k)] = 0.0;
        }
    for (int i = 0; i < g.Nx; i++)
        for (int k = 0; k < g.Nz; k++) {
            if (!outflow(i, 0, k)) g.v[g.v_idx(i, 0, k)] = 0.0;
            if (!outflow(i, g.Ny - 1, k)) g.v[g.v_idx(i, g.Ny, k)] = 0.0;
        }
    for (int i = 0; i < g.Nx; i++)
        for (int j = 0; j < g.Ny; j++) {
            if (!outflow(i, j, 0)) g.w[g.w_idx(i, j, 0)] = 0.0;
            if (!outflow(i, j, g.Nz - 1)) g.w[g.w_idx(i, j, g.Nz)] = 0.0;
        }
    for (int i = 0; i < g.Nx; i++)
        for (int j = 0; j < g.Ny; j++)
            for (int k = 0; k < g.Nz; k++) {
                if (g.cell_type[g.c_idx(i, j, k)] != CellType::SOLID) continue;
                g.u[g.u_idx(i, j, k)] = 0.0; g.u[g.u_idx(i + 1, j, k)] = 0.0;
                g.v[g.v_idx(i, j, k)] = 0.0; g.v[g.v_idx(i, j + 1, k)] = 0.0;
                g.w[g.w_idx(i, j, k)] = 0.0; g.w[g.w_idx(i, j, k + 1)] = 0.0;
            }
}

void BoundaryManager::apply_inflow(Grid &g)
{   // :111-130
    HostEdit he(g);
    for (const auto &f : fans_)
        for (int i = std::max(0, f.i_min); i <= std::min(g.Nx - 1, f.i_max); i++)
            for (int j = std::max(0, f.j_min); j <= std::min(g.Ny - 1, f.j_max); j++)
                for (int k = std::max(0, f.k_min); k <= std::min(g.Nz - 1, f.k_max); k++) {
                    if (f.vel[0] != 0.0) { g.u[g.u_idx(i, j, k)] = f.vel[0]; g.u[g.u_idx(i + 1, j, k)] = f.vel[0]; }
                    if (f.vel[1] != 0.0) { g.v[g.v_idx(i, j, k)] = f.vel[1]; g.v[g.v_idx(i, j + 1, k)] = f.vel[1]; }
                    if (f.vel[2] != 0.0) { g.w[g.w_idx(i, j, k)] = f.vel[2]; g.w[g.w_idx(i, j, k + 1)] = f.vel[2]; }
                }
}

void BoundaryManager::apply_outflow(Grid &g)
{   // :180-220
    HostEdit he(g);
    for (const auto &o : outflows_)
        for (int a = o.a_min; a <= o.a_max; a++)
            for (int b = o.b_min; b <= o.b_max; b++) {
                if (o.axis == 0) {
                    if (o.side) g.u[g.u_idx(g.Nx, a, b)] = g.u[g.u_idx(g.Nx - 1, a, b)];
                    else g.u[g.u_idx(0, a, b)] = g.u[g.u_idx(1, a, b)];
                } else if (o.axis == 1) {
                    if (o.side) g.v[g.v_idx(a, g.Ny, b)] = g.v[g.v_idx(a, g.Ny - 1, b)];
                    else g.v[g.v_idx(a, 0, b)] = g.v[g.v_idx(a, 1, b)];
                } else {
                    if (o.side) g.w[g.w_idx(a, b, g.Nz)] = g.w[g.w_idx(a, b, g.Nz - 1)];
                    else g.w[g.w_idx(a, b, 0)] = g.w[g.w_idx(a, b, 1)];
                }
            }
}

void BoundaryManager::apply_opening_temperature(Grid &g)
{   // :137-178
    HostEdit he(g);
    for (const auto &o : openings_) {
        const int sd = o.side;
        for (int a = o.a_min; a <= o.a_max; a++)
            for (int b = o.b_min; b <= o.b_max; b++) {
                double vf;
                int wall, inner;
                if (o.axis == 0) {
                    vf = g.u[g.u_idx(sd ? g.Nx : 0, a, b)];
                    wall = g.c_idx(sd ? g.Nx - 1 : 0, a, b);
                    inner = g.c_idx(sd ? g.Nx - 2 : 1, a, b);
                } else if (o.axis == 1) {
                    vf = g.v[g.v_idx(a, sd ? g.Ny : 0, b)];
                    wall = g.c_idx(a, sd ? g.Ny - 1 : 0, b);
                    inner = g.c_idx(a, sd ? g.Ny - 2 : 1, b);
                } else {
                    vf = g.w[g.w_idx(a, b, sd ? g.Nz : 0)];
                    wall = g.c_idx(a, b, sd ? g.Nz - 1 : 0);
                    inner = g.c_idx(a, b, sd ? g.Nz - 2 : 1);
                }
                const bool incoming = sd ? (vf <= 0.0) : (vf >= 0.0);
                g.T[wall] = incoming ? o.T_outside : g.T[inner];
            }
    }
}

}  // namespace ffroom
