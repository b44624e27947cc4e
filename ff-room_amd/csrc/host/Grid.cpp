// Grid.cpp - host MAC grid (reference: cpp/src/core/Grid.cpp:6-57).
#include "Grid.hpp"

#include <algorithm>
#include <stdexcept>

#include "FluidSolver.hpp"

namespace ffroom {

Grid::Grid(int nx, int ny, int nz, double lx, double ly, double lz)
    : Nx(nx), Ny(ny), Nz(nz), dx(lx / nx), dy(ly / ny), dz(lz / nz)
{
    if (nx <= 0 || ny <= 0 || nz <= 0) throw std::invalid_argument("Grid dimensions must be positive");
    const size_t nu = (size_t)(nx + 1) * ny * nz, nv = (size_t)nx * (ny + 1) * nz;
    const size_t nw = (size_t)nx * ny * (nz + 1), nc = (size_t)nx * ny * nz;
    u.assign(nu, 0.0); v.assign(nv, 0.0); w.assign(nw, 0.0); p.assign(nc, 0.0);
    u_tmp.assign(nu, 0.0); v_tmp.assign(nv, 0.0); w_tmp.assign(nw, 0.0);
    cell_type.assign(nc, CellType::FLUID);
    T.assign(nc, 0.0); T_tmp.assign(nc, 0.0);
    xs.resize(nx + 1); ys.resize(ny + 1); zs.resize(nz + 1);
    for (int i = 0; i <= nx; i++) xs[i] = i * dx;
    for (int j = 0; j <= ny; j++) ys[j] = j * dy;
    for (int k = 0; k <= nz; k++) zs[k] = k * dz;
    dxv.assign(nx, dx); dyv.assign(ny, dy); dzv.assign(nz, dz);
}

void Grid::set_face_coords(std::vector<double> xs_, std::vector<double> ys_, std::vector<double> zs_)
{
    if ((int)xs_.size() != Nx + 1 || (int)ys_.size() != Ny + 1 || (int)zs_.size() != Nz + 1)
        throw std::invalid_argument("face coordinate arrays must have N+1 entries");
    xs = std::move(xs_); ys = std::move(ys_); zs = std::move(zs_);
    for (int i = 0; i < Nx; i++) dxv[i] = xs[i + 1] - xs[i];
    for (int j = 0; j < Ny; j++) dyv[j] = ys[j + 1] - ys[j];
    for (int k = 0; k < Nz; k++) dzv[k] = zs[k + 1] - zs[k];
    geometry_version++;
}

void Grid::zero_velocity()
{
    sync_to_host();
    std::fill(u.begin(), u.end(), 0.0);
    std::fill(v.begin(), v.end(), 0.0);
    std::fill(w.begin(), w.end(), 0.0);
    mark_host_modified();
}

void Grid::zero_pressure()
{
    sync_to_host();
    std::fill(p.begin(), p.end(), 0.0);
    mark_host_modified();
}

void Grid::sync_to_host()
{
    if (attached) attached->sync_to_host();
}

void Grid::mark_host_modified()
{
    if (attached) attached->mark_host_modified();
}

}  // namespace ffroom
