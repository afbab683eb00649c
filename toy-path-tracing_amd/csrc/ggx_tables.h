// GGX tables read by the OpenPBR BSDF on the device, built on the host the
// way the reference builds them (R/bsdf/directional_albedo.rs):
//   directional albedo of a rough dielectric (reflection only), per IOR,
//     64 sqrt(cos) x 16 phi x 64 roughness x 2 anisotropy slices
//   conductor energy compensation, white Fresnel: e 64 roughness x 64 sqrt(cos), e_avg 64
//   dielectric energy compensation, full sphere: e 32 eta x 32 roughness x 32 sqrt(cos),
//     e_avg 32 eta x 32 roughness
// Process-wide cache: each table is built once.
#pragma once

#include <cstdint>
#include <vector>

namespace tpt {
namespace ggx {

constexpr uint32_t kAlbedoCos = 64, kAlbedoPhi = 16, kAlbedoRough = 64, kAlbedoAniso = 2;
constexpr uint32_t kAlbedoLen = kAlbedoCos * kAlbedoPhi * kAlbedoRough * kAlbedoAniso;
constexpr uint32_t kCondCos = 64, kCondRough = 64;
constexpr uint32_t kDielCos = 32, kDielRough = 32, kDielEta = 32;

// The table of one IOR (sanitised: non-finite or non-positive -> 1.5).
const std::vector<float>& albedo_table(float eta);
// e (kCondRough * kCondCos) then e_avg (kCondRough).
const std::vector<float>& conductor_table();
// e (kDielEta * kDielRough * kDielCos) then e_avg (kDielEta * kDielRough).
const std::vector<float>& dielectric_table();

// OpenPBR's specular IOR (modulated by the specular weight,
// R/material/open_pbr.rs:871-878), the key of a material's albedo table.
float modulated_eta(float eta, float weight);

}  // namespace ggx
}  // namespace tpt
