// Sphere tracing of a device-resident scene group: the reference's
// SphereTracer::Trace (util/sphere_tracer.hpp:46-268) with its PinholeCamera
// (geometry/pinhole_camera.hpp:9-66), in double precision on the device.
#pragma once

#include <cstdint>

#include "../../include/prtcl_hip.h"
#include "scene_impl.hpp"

namespace prtcl_amd {

// traces the particles of group `group` (the tracer's visible group) seen
// by `cam`; image and parameter (may be null) are host arrays of
// width * height doubles, row iy at [iy * width + ix]
void trace_scene(SceneBase &sc, CtxResources &res, int group, const prtcl_hip_camera &cam,
                 double threshold, int32_t max_steps, double *image, double *parameter);

}  // namespace prtcl_amd
