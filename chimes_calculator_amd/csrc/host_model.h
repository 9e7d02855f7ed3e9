// Host-side bridge from a parsed parameter set to the engine: device model
// construction and the per-device engine registry.
#pragma once

#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "chx_engine.h"
#include "params.h"

namespace chx {

// One HIP engine per device, created on first use and kept for the process.
struct EngineHandle {
    chx_engine *eng = nullptr;
    std::mutex mu;
};
EngineHandle *engine_for_device(int device, std::string &err);
int default_device();

// A parameter set resident on one device.
struct DeviceModel {
    ParamSet P;
    EngineHandle *eh = nullptr;
    chx_model *m = nullptr;
    int device = 0;
    double rc[3] = {0, 0, 0};
    ~DeviceModel();
};

// Builds the dense tables and uploads them.  Returns nullptr with err set on
// failure.
std::unique_ptr<DeviceModel> make_device_model(const ParamSet &P, int device, std::string &err);

}  // namespace chx

// The public handle of chimes_gpu.h.
struct chimes_gpu_model {
    std::unique_ptr<chx::DeviceModel> dm;
};
