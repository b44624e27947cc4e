// Device.cpp - see Device.hpp.
#include "Device.hpp"

#include <cstdlib>
#include <mutex>
#include <vector>

namespace ffroom {

namespace {
std::mutex g_mu;
ffr_ctx *g_ctx = nullptr;
}  // namespace

ffr_ctx *device_context()
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ctx) {
        int dev = 0;
        if (const char *e = std::getenv("FFROOM_DEVICE")) dev = std::atoi(e);
        int rc = ffr_ctx_create(dev, 0, &g_ctx);
        if (rc != FFR_OK) {
            g_ctx = nullptr;
            throw std::runtime_error(std::string("ff-room GPU context: ") + ffr_last_error(nullptr) +
                                     " (" + ffr_status_string(rc) + ")");
        }
    }
    return g_ctx;
}

void ffr_check(int rc, const char *what)
{
    if (rc == FFR_OK) return;
    const char *msg = g_ctx ? ffr_last_error(g_ctx) : ffr_last_error(nullptr);
    throw std::runtime_error(std::string(what) + ": " + ffr_status_string(rc) + ": " + msg);
}

DeviceScene::DeviceScene(const Grid &g, const BoundaryManager *bm, const ffr_params &prm)
{
    ffr_ctx *ctx = device_context();
    std::vector<uint8_t> flags(g.cell_type.size());
    for (size_t i = 0; i < flags.size(); i++) flags[i] = static_cast<uint8_t>(g.cell_type[i]);
    std::vector<ffr_fan> fans;
    std::vector<ffr_outflow> outs;
    std::vector<ffr_opening> opens;
    if (bm) {
        for (auto &f : bm->fans())
            fans.push_back({f.i_min, f.i_max, f.j_min, f.j_max, f.k_min, f.k_max,
                            {f.vel[0], f.vel[1], f.vel[2]}});
        for (auto &o : bm->outflows())
            outs.push_back({o.axis, o.side, o.a_min, o.a_max, o.b_min, o.b_max});
        for (auto &o : bm->openings())
            opens.push_back({o.axis, o.side, o.a_min, o.a_max, o.b_min, o.b_max, o.T_outside});
    }
    ffr_scene_desc d{};
    d.nx = g.Nx; d.ny = g.Ny; d.nz = g.Nz;
    d.dxv = g.dxv.data(); d.dyv = g.dyv.data(); d.dzv = g.dzv.data();
    d.cell_type = flags.data();
    d.fans = fans.data(); d.n_fans = (int)fans.size();
    d.outflows = outs.data(); d.n_outflows = (int)outs.size();
    d.openings = opens.data(); d.n_openings = (int)opens.size();
    ffr_check(ffr_scene_create(ctx, &d, &prm, &scene_), "ffr_scene_create");
    version_ = g.geometry_version + (bm ? bm->version() : 0);
}

DeviceScene::~DeviceScene()
{
    if (scene_) ffr_scene_destroy(scene_);
}

}  // namespace ffroom
