// Model archives of device-resident scenes: checkpoint and resume in the
// reference's native binary model format (util/archive.cpp:4-83 for sizes,
// strings and values; data/model.cpp:54-75, group.cpp:31-59,
// varying_manager.cpp:21-45, uniform_manager.cpp:22-41 for the layout).
//
// A scene is written as a model with the global fields its schemes bind, a
// group "b" of type "boundary" and a group "f" of type "fluid" (tag
// "dynamic"; the names of the reference's scripts, example.lua:106-110),
// each with the varying fields the scene holds, in host order, and its
// uniform fields. Three extra fields carry the device state a bit-identical
// continuation needs and the reference loads as ordinary fields:
//   global "device_clock" f64 (the step clock; current_time is rounded),
//   global "device_steps_since_permute" s32,
//   varying "device_slot" s32 in each group (the slot of every particle).
#pragma once

#include <cstdint>
#include <memory>

#include "../../include/prtcl_hip.h"
#include "scene_impl.hpp"

namespace prtcl_amd {

// writes the scene to `path`; returns the bytes written
uint64_t save_scene_archive(SceneBase &sc, CtxResources &res, const char *path);

struct ParsedArchive;
struct ParsedArchiveDeleter {
  void operator()(ParsedArchive *p) const;
};
using ParsedArchivePtr = std::unique_ptr<ParsedArchive, ParsedArchiveDeleter>;

// reads the structure of the archive at `path` (values of the varying
// fields stay in the file) and completes `desc`: particle counts, real type,
// the model's parameters (globals smoothing_scale, time_step, fade_duration,
// iisph_relaxation, gravity; fluid uniforms rest_density, dynamic_viscosity,
// wkbb18_maximum_error / _iterations, strain_rate_viscosity; boundary
// uniform dynamic_viscosity) and the implicit viscosity scheme its fluid
// fields identify. Throws std::invalid_argument on a malformed archive.
ParsedArchivePtr parse_scene_archive(const char *path, prtcl_hip_scene_desc &desc);

// fills a scene created from the completed description: slots first, then
// every varying field the scene holds, then the clock and solver state
uint64_t load_scene_archive(SceneBase &sc, const ParsedArchive &a);

// field name of a C ABI field id (capi.hip)
const char *field_name(int id);

}  // namespace prtcl_amd
