// module.cpp - the `_ffroom_core` extension module, with every name and
// signature of the reference's bindings (cpp/src/bindings/module.cpp:15-210),
// LBMSolver included. Field views are zero-copy and writable as in the
// reference. The state lives on the GPU, so taking a view first pulls the
// current fields into the Grid and marks them modified; while any view is
// alive, every step leaves the Grid vectors current and uploads them again
// before the next step, so a view held across steps shows the latest fields
// and edits made through it reach the GPU, as the reference's live vectors do.
#include <pybind11/functional.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cstring>

#include "../host/BoundaryManager.hpp"
#include "../host/Device.hpp"
#include "../host/FluidSolver.hpp"
#include "../host/Grid.hpp"
#include "../host/LBMSolver.hpp"
#include "../host/PoissonSolver.hpp"

namespace py = pybind11;
using namespace ffroom;

namespace {
// Base object of a field view: keeps the Grid alive and counts the view in
// Grid::live_views until the array is released.
struct ViewOwner {
    py::object grid;
    Grid *g;
};

template <typename T>
py::array_t<double> view(Grid &g, std::vector<double> &vec, std::vector<ssize_t> shape)
{
    g.sync_to_host();
    g.mark_host_modified();
    auto *owner = new ViewOwner{py::cast(g), &g};
    g.live_views++;
    py::capsule base(owner, [](void *p) {
        auto *o = static_cast<ViewOwner *>(p);
        o->g->live_views--;
        delete o;
    });
    return py::array_t<double>(shape, vec.data(), base);
}
}  // namespace

PYBIND11_MODULE(_ffroom_core, m)
{
    m.doc() = "ff-room CFD core (projection method on MAC grid), MI355X GPU build";

    py::enum_<AdvectionScheme>(m, "AdvectionScheme")
        .value("UPWIND1", AdvectionScheme::UPWIND1)
        .value("QUICK", AdvectionScheme::QUICK)
        .value("LAX_WENDROFF", AdvectionScheme::LAX_WENDROFF)
        .export_values();

    py::enum_<CellType>(m, "CellType")
        .value("FLUID", CellType::FLUID)
        .value("SOLID", CellType::SOLID)
        .value("INFLOW", CellType::INFLOW)
        .value("OUTFLOW", CellType::OUTFLOW)
        .export_values();

    py::enum_<Precision>(m, "Precision")
        .value("FP64", Precision::FP64)
        .value("MIXED", Precision::MIXED);

    py::enum_<PoissonMethod>(m, "PoissonMethod")
        .value("AUTO", PoissonMethod::AUTO)
        .value("CG", PoissonMethod::CG)
        .value("MGPCG", PoissonMethod::MGPCG);

    py::class_<Grid>(m, "Grid")
        .def(py::init<int, int, int, double, double, double>(), py::arg("nx"), py::arg("ny"),
             py::arg("nz"), py::arg("lx"), py::arg("ly"), py::arg("lz"))
        .def_readonly("Nx", &Grid::Nx)
        .def_readonly("Ny", &Grid::Ny)
        .def_readonly("Nz", &Grid::Nz)
        .def_readonly("dx", &Grid::dx)
        .def_readonly("dy", &Grid::dy)
        .def_readonly("dz", &Grid::dz)
        .def_property_readonly("u", [](Grid &g) { return view<double>(g, g.u, {(ssize_t)g.u.size()}); })
        .def_property_readonly("v", [](Grid &g) { return view<double>(g, g.v, {(ssize_t)g.v.size()}); })
        .def_property_readonly("w", [](Grid &g) { return view<double>(g, g.w, {(ssize_t)g.w.size()}); })
        .def_property_readonly("p", [](Grid &g) { return view<double>(g, g.p, {(ssize_t)g.p.size()}); })
        .def("u_shaped", [](Grid &g) { return view<double>(g, g.u, {g.Nx + 1, g.Ny, g.Nz}); })
        .def("v_shaped", [](Grid &g) { return view<double>(g, g.v, {g.Nx, g.Ny + 1, g.Nz}); })
        .def("w_shaped", [](Grid &g) { return view<double>(g, g.w, {g.Nx, g.Ny, g.Nz + 1}); })
        .def("p_shaped", [](Grid &g) { return view<double>(g, g.p, {g.Nx, g.Ny, g.Nz}); })
        .def("T_shaped", [](Grid &g) { return view<double>(g, g.T, {g.Nx, g.Ny, g.Nz}); })
        .def("cell_type_shaped",
             [](Grid &g) {
                 auto arr = py::array_t<uint8_t>({g.Nx, g.Ny, g.Nz});
                 std::memcpy(arr.mutable_data(), g.cell_type.data(), g.cell_type.size());
                 return arr;
             })
        .def("xs", [](Grid &g) { return py::array_t<double>({g.Nx + 1}, g.xs.data(), py::cast(g)); })
        .def("ys", [](Grid &g) { return py::array_t<double>({g.Ny + 1}, g.ys.data(), py::cast(g)); })
        .def("zs", [](Grid &g) { return py::array_t<double>({g.Nz + 1}, g.zs.data(), py::cast(g)); })
        .def("set_face_coords", &Grid::set_face_coords, py::arg("xs"), py::arg("ys"), py::arg("zs"))
        .def("zero_velocity", &Grid::zero_velocity)
        .def("zero_pressure", &Grid::zero_pressure);

    py::class_<FanBC>(m, "FanBC")
        .def(py::init<>())
        .def_readwrite("i_min", &FanBC::i_min)
        .def_readwrite("i_max", &FanBC::i_max)
        .def_readwrite("j_min", &FanBC::j_min)
        .def_readwrite("j_max", &FanBC::j_max)
        .def_readwrite("k_min", &FanBC::k_min)
        .def_readwrite("k_max", &FanBC::k_max)
        .def_readwrite("vel", &FanBC::vel);

    py::class_<OutflowBC>(m, "OutflowBC")
        .def(py::init<>())
        .def_readwrite("axis", &OutflowBC::axis)
        .def_readwrite("side", &OutflowBC::side)
        .def_readwrite("a_min", &OutflowBC::a_min)
        .def_readwrite("a_max", &OutflowBC::a_max)
        .def_readwrite("b_min", &OutflowBC::b_min)
        .def_readwrite("b_max", &OutflowBC::b_max);

    py::class_<OpeningBC, OutflowBC>(m, "OpeningBC")
        .def(py::init<>())
        .def_readwrite("axis", &OpeningBC::axis)
        .def_readwrite("side", &OpeningBC::side)
        .def_readwrite("a_min", &OpeningBC::a_min)
        .def_readwrite("a_max", &OpeningBC::a_max)
        .def_readwrite("b_min", &OpeningBC::b_min)
        .def_readwrite("b_max", &OpeningBC::b_max)
        .def_readwrite("T_outside", &OpeningBC::T_outside);

    py::class_<BoundaryManager>(m, "BoundaryManager")
        .def(py::init<>())
        .def("add_solid_box", &BoundaryManager::add_solid_box)
        .def("add_fan", &BoundaryManager::add_fan)
        .def("add_outflow", &BoundaryManager::add_outflow)
        .def("add_opening", &BoundaryManager::add_opening)
        .def("apply_noslip", &BoundaryManager::apply_noslip)
        .def("apply_inflow", &BoundaryManager::apply_inflow)
        .def("apply_outflow", &BoundaryManager::apply_outflow)
        .def("apply_opening_temperature", &BoundaryManager::apply_opening_temperature)
        // extensions
        .def("set_fan_temperature", &BoundaryManager::set_fan_temperature, py::arg("index"),
             py::arg("T_inlet"))
        .def("fans", &BoundaryManager::fans)
        .def("outflows", &BoundaryManager::outflows)
        .def("openings", &BoundaryManager::openings);

    py::class_<PoissonSolverParams>(m, "PoissonSolverParams")
        .def(py::init<>())
        .def_readwrite("max_iter", &PoissonSolverParams::max_iter)
        .def_readwrite("tol", &PoissonSolverParams::tol);

    py::class_<FluidSolverParams>(m, "FluidSolverParams")
        .def(py::init<>())
        .def_readwrite("rho", &FluidSolverParams::rho)
        .def_readwrite("nu", &FluidSolverParams::nu)
        .def_readwrite("dt", &FluidSolverParams::dt)
        .def_readwrite("max_steps", &FluidSolverParams::max_steps)
        .def_readwrite("convergence_tol", &FluidSolverParams::convergence_tol)
        .def_readwrite("poisson", &FluidSolverParams::poisson)
        .def_readwrite("advection", &FluidSolverParams::advection)
        .def_readwrite("use_openmp", &FluidSolverParams::use_openmp)
        .def_readwrite("use_multigrid", &FluidSolverParams::use_multigrid)
        .def_readwrite("thermal", &FluidSolverParams::thermal)
        .def_readwrite("T_initial", &FluidSolverParams::T_initial)
        .def_readwrite("T_target", &FluidSolverParams::T_target)
        .def_readwrite("buoyancy", &FluidSolverParams::buoyancy)
        .def_readwrite("g_accel", &FluidSolverParams::g_accel)
        .def_readwrite("beta", &FluidSolverParams::beta)
        .def_readwrite("k_thermal", &FluidSolverParams::k_thermal)
        .def_readwrite("cp", &FluidSolverParams::cp)
        // extensions
        .def_readwrite("precision", &FluidSolverParams::precision)
        .def_readwrite("poisson_method", &FluidSolverParams::poisson_method)
        .def_readwrite("mg_levels", &FluidSolverParams::mg_levels)
        .def_readwrite("mg_bf16", &FluidSolverParams::mg_bf16)
        .def_readwrite("inner_tol", &FluidSolverParams::inner_tol)
        .def_readwrite("inner_max_iter", &FluidSolverParams::inner_max_iter)
        .def_readwrite("sync_host_every_step", &FluidSolverParams::sync_host_every_step);

    py::class_<StepResult>(m, "StepResult")
        .def_readonly("step", &StepResult::step)
        .def_readonly("velocity_change", &StepResult::velocity_change)
        .def_readonly("pressure_residual", &StepResult::pressure_residual)
        .def_readonly("divergence_max", &StepResult::divergence_max)
        .def_readonly("converged", &StepResult::converged)
        .def_readonly("T_mean", &StepResult::T_mean)
        .def_readonly("poisson_iterations", &StepResult::poisson_iterations)
        .def_readonly("nonfinite", &StepResult::nonfinite)
        .def("__repr__", [](const StepResult &r) {
            return "StepResult(step=" + std::to_string(r.step) +
                   ", vel_change=" + std::to_string(r.velocity_change) +
                   ", p_res=" + std::to_string(r.pressure_residual) +
                   ", div_max=" + std::to_string(r.divergence_max) +
                   ", T_mean=" + std::to_string(r.T_mean) +
                   ", converged=" + (r.converged ? "True" : "False") + ")";
        });

    py::class_<FluidSolver>(m, "FluidSolver")
        .def(py::init<Grid &, BoundaryManager &, FluidSolverParams>(), py::arg("grid"),
             py::arg("bm"), py::arg("params") = FluidSolverParams{})
        .def("step", &FluidSolver::step, py::call_guard<py::gil_scoped_release>())
        .def("run", &FluidSolver::run, py::arg("callback") = nullptr,
             py::call_guard<py::gil_scoped_release>())
        .def("compute_divergence_max", &FluidSolver::compute_divergence_max)
        .def_property_readonly("current_step", &FluidSolver::current_step)
        // extensions
        .def("compute_T_mean", &FluidSolver::compute_T_mean)
        .def("sync_to_host", &FluidSolver::sync_to_host)
        .def(
            "snapshot_begin",
            [](FluidSolver &s, bool cells) { return s.snapshot_begin(cells ? FFR_SNAP_CELLS : FFR_SNAP_FIELDS); },
            py::arg("cells") = false)
        .def(
            "snapshot_end",
            // fields: (step, u, v, w, p, T); cells: (step, vel, speed, T)
            [](FluidSolver &s, int slot) -> py::tuple {
                const Grid &g = s.grid();
                int step;
                if (s.snapshot_kind(slot) == FFR_SNAP_CELLS) {
                    py::array_t<double> vel({g.Nx, g.Ny, g.Nz, 3}), speed({g.Nx, g.Ny, g.Nz}),
                        T({g.Nx, g.Ny, g.Nz});
                    double *pv = vel.mutable_data(), *ps = speed.mutable_data(), *pT = T.mutable_data();
                    {
                        py::gil_scoped_release nogil;
                        step = s.snapshot_end(slot, pv, ps, pT, nullptr, nullptr);
                    }
                    return py::make_tuple(step, vel, speed, T);
                }
                py::array_t<double> u({g.Nx + 1, g.Ny, g.Nz}), v({g.Nx, g.Ny + 1, g.Nz}),
                    w({g.Nx, g.Ny, g.Nz + 1}), p({g.Nx, g.Ny, g.Nz}), T({g.Nx, g.Ny, g.Nz});
                double *pu = u.mutable_data(), *pv = v.mutable_data(), *pw = w.mutable_data(),
                       *pp = p.mutable_data(), *pT = T.mutable_data();
                {
                    py::gil_scoped_release nogil;
                    step = s.snapshot_end(slot, pu, pv, pw, pp, pT);
                }
                return py::make_tuple(step, u, v, w, p, T);
            },
            py::arg("slot"));

    // LBMSolver (module.cpp:202-209 of the reference): the D3Q19 lattice on the GPU
    py::class_<LBMSolver>(m, "LBMSolver")
        .def(py::init<Grid &, BoundaryManager &, double, int, double, double>(), py::arg("grid"),
             py::arg("bm"), py::arg("tau"), py::arg("max_steps"), py::arg("convergence_tol"),
             py::arg("dt") = 0.01)
        .def("step", &LBMSolver::step, py::call_guard<py::gil_scoped_release>())
        .def("run", &LBMSolver::run, py::arg("callback") = nullptr,
             py::call_guard<py::gil_scoped_release>())
        .def("compute_divergence_max", &LBMSolver::compute_divergence_max)
        .def_property_readonly("current_step", &LBMSolver::current_step)
        .def("compute_T_mean", &LBMSolver::compute_T_mean)
        .def("sync_to_host", &LBMSolver::sync_to_host);

    // extensions: the pressure solvers on their own
    py::class_<PoissonSolver>(m, "PoissonSolver")
        .def(py::init<PoissonSolverParams>(), py::arg("params") = PoissonSolverParams{})
        .def("solve", &PoissonSolver::solve, py::arg("grid"), py::arg("rhs"))
        .def("last_iterations", &PoissonSolver::last_iterations)
        .def("last_residual", &PoissonSolver::last_residual);
    py::class_<MultigridPoissonSolver, PoissonSolver>(m, "MultigridPoissonSolver")
        .def(py::init<PoissonSolverParams>(), py::arg("params") = PoissonSolverParams{});

    m.attr("ABI_VERSION") = ffr_abi_version();
    m.attr("HAS_LBM") = true;
}
