// Grid.hpp - host-side MAC grid, same public interface as the reference's
// cpp/src/core/Grid.hpp:26-69 (public std::vector fields, k-fastest flat
// indexing). The simulation state itself lives on the GPU while a solver
// (FluidSolver or LBMSolver) is attached; the vectors here are its host
// mirror, refreshed by the solver's sync_to_host() (called at the end of
// run() and by the Python views) and pushed back by mark_host_modified().
#pragma once
#include <array>
#include <cstdint>
#include <vector>

namespace ffroom {

enum class CellType : uint8_t { FLUID = 0, SOLID = 1, INFLOW = 2, OUTFLOW = 3 };

// A solver that keeps a Grid's state on the GPU.
struct DeviceMirror {
    virtual void sync_to_host() = 0;        // pull the state if the device is ahead
    virtual void mark_host_modified() = 0;  // push host edits before the next step

protected:
    ~DeviceMirror() = default;
};

struct Grid {
    int Nx, Ny, Nz;
    double dx, dy, dz;

    std::vector<double> dxv, dyv, dzv;  // cell widths
    std::vector<double> xs, ys, zs;     // face coordinates

    void set_face_coords(std::vector<double> xs_, std::vector<double> ys_, std::vector<double> zs_);

    std::vector<double> u, v, w, p;
    std::vector<double> u_tmp, v_tmp, w_tmp;
    std::vector<CellType> cell_type;
    std::vector<double> T, T_tmp;

    Grid(int nx, int ny, int nz, double lx, double ly, double lz);

    inline int u_idx(int i, int j, int k) const { return i * (Ny * Nz) + j * Nz + k; }
    inline int v_idx(int i, int j, int k) const { return i * ((Ny + 1) * Nz) + j * Nz + k; }
    inline int w_idx(int i, int j, int k) const { return i * (Ny * (Nz + 1)) + j * (Nz + 1) + k; }
    inline int c_idx(int i, int j, int k) const { return i * (Ny * Nz) + j * Nz + k; }

    void zero_velocity();
    void zero_pressure();

    // ---- device mirror bookkeeping (no counterpart in the reference) ------
    // Bumped whenever flags, spacings or boundary lists change, so that an
    // attached solver rebuilds its device geometry before the next step.
    uint64_t geometry_version = 0;
    DeviceMirror *attached = nullptr;
    // Field views (the Python u/v/w/p/T arrays over these vectors) still
    // alive: while there are any, a step leaves the vectors current and
    // uploads them again before the next step, so a held view behaves as the
    // reference's view of its live vector.
    int live_views = 0;
    // Makes the host vectors current (pulls from the GPU if needed).
    void sync_to_host();
    // Declares that host vectors were (or may be) written: the next step
    // pushes them to the GPU first.
    void mark_host_modified();
};

}  // namespace ffroom
