// Legacy-VTK output of a device-resident scene group, byte for byte the text
// of the reference's SaveVTK (util/save_vtk.cpp:57-129): an ASCII POLYDATA
// file with the group's positions as POINTS and its velocity, acceleration,
// host index, mass, volume, density, pressure and time_of_birth as
// POINT_DATA, every real printed as std::fixed with precision 6 (the
// "%.6f" of the C library: the exact binary value rounded half to even).
// The numbers are formatted on the device.
#pragma once

#include <cstdint>

#include "scene_impl.hpp"

namespace prtcl_amd {

// writes group `group` of the scene to `path`; returns the bytes written
uint64_t save_scene_vtk(SceneBase &sc, CtxResources &res, int group, const char *path);

}  // namespace prtcl_amd
