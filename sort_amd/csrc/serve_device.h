// Factories of the serving engine's devices.
#pragma once
#include <memory>
#include <string>

#include "api_internal.h"
#include "serve_engine.h"

namespace sortgpu {
namespace serve {

// The HIP device: `streams` non-blocking streams on the context's device, each with a wave
// workspace of `wave_paths` path slots, and scene and tree tables of `max_scenes` entries. With
// `graphs`, each wave's bounce loop is replayed from a captured HIP graph.
int make_hip_device(sortgpu_ctx* ctx, uint32_t wave_paths, uint32_t max_scenes, uint32_t streams, bool graphs,
                    std::unique_ptr<Device>& out, std::string& err);

// The host-only stand-in. A wave reports done after `polls` non-blocking queries.
std::unique_ptr<Device> make_host_device(sortgpu_ctx* ctx, uint32_t wave_paths, uint32_t polls);

}  // namespace serve
}  // namespace sortgpu
