// Device.hpp - the host classes' handle on the C ABI: one process-wide
// context per GPU and a scene built from a Grid + BoundaryManager.
#pragma once
#include <stdexcept>
#include <string>

#include "../../../include/ffroom.h"
#include "BoundaryManager.hpp"
#include "Grid.hpp"

namespace ffroom {

// Process-wide context on device FFROOM_DEVICE (default 0), created on first use.
ffr_ctx *device_context();

// Throws std::runtime_error carrying ffr_last_error() when rc != FFR_OK.
void ffr_check(int rc, const char *what);

class DeviceScene {
public:
    // bm may be null (a Poisson-only scene: flags and spacings only).
    DeviceScene(const Grid &grid, const BoundaryManager *bm, const ffr_params &prm);
    ~DeviceScene();
    DeviceScene(const DeviceScene &) = delete;
    DeviceScene &operator=(const DeviceScene &) = delete;
    ffr_scene *get() const { return scene_; }
    uint64_t geometry_version() const { return version_; }

private:
    ffr_scene *scene_ = nullptr;
    uint64_t version_ = 0;
};

}  // namespace ffroom
