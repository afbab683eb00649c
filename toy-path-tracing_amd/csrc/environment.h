// Environment light: host build (environment.cpp) of the data the device
// samples.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace tpt {

struct HostEnv {
  uint32_t w = 0, h = 0;
  float scale = 1.0f;
  std::vector<float> px;        // RGB, row-major, 3 * w * h
  std::vector<float> pyr;       // MIS-compensated weight pyramid, levels coarsest first
  std::vector<uint32_t> levels; // per level: offset into pyr, width, height
  uint32_t level_count = 0;
  float total_plain = 0.0f;     // total of the plain weights (total_power)
  float total_mis = 0.0f;       // total of the MIS-compensated weights
  float rot[9], inv_rot[9];     // Mat3::from_rotation_y(+-rotate_y), columns
  float power = 0.0f;           // EnvironmentLight::total_power
};

bool build_environment(const float* px, uint32_t w, uint32_t h, float scale, float rotate_y, HostEnv* e,
                       std::string* err);

}  // namespace tpt
