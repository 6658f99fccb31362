// stt_host.hpp — host-side helpers of the library (not part of the C-ABI).
#pragma once
#include <string>

#include "../../include/stt.h"
#include "stt_device.hpp"

namespace stt {

// Records msg as this thread's stt_last_error() text and returns code.
int set_error(int code, const std::string& msg);

// Quaternion (x, y, z, w) of a rotation by theta degrees about +y, as the
// reference's InitViews builds it.
void quat_about_y(float theta_deg, float q[4]);
// Rotation by deg degrees about +y composed with q_old (RotateVolume).
void quat_rotate_y(const float q_old[4], double deg, float q[4]);
// The Scene's modelview and projection matrices (column-major) of a rotation.
void scene_matrices(const stt_geometry& g, const float q[4], float view[16], float proj[16]);
// Per-view geometry record from a modelview and a projection matrix, and
// from a rotation (fp32, reference arithmetic order).
ViewRec build_view_from(const stt_geometry& g, const float view[16], const float proj[16]);
ViewRec build_view(const stt_geometry& g, const float q[4]);
// Sample distance ||spacing||.
float sample_distance(const stt_geometry& g);

}  // namespace stt
