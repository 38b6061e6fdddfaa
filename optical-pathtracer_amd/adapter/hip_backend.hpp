// The hip backend of the Optical World Engine: owe::Backend and owe::Renderer (render/backend.hpp,
// render/renderer.hpp) over the C ABI of include/owe_hip.h. Maintainer-side source: it is compiled
// inside the engine as src/owe/backends/hip/hip_backend.{hpp,cpp} and linked with libowe_hip.so.
// Like the gpu backend (backends/gpu/gpu_backend.hpp) it depends only on the render interface and the
// scene model, and carries no HIP types.
#pragma once

#include <string>
#include <vector>

#include "owe/render/backend.hpp"
#include "owe_hip_scene.h"

namespace owe {

const Backend& hipBackend();

// The transport-relevant part of a built scene as the plain structs of owe_hip_scene.h, for one
// detector. The arrays live in this object; `scene` points into them.
struct HipFlatScene {
    std::vector<owe_medium> media;
    std::vector<uint32_t> regionMedium;
    std::vector<owe_optics> optics;
    std::vector<owe_emission> emissions;
    std::vector<owe_boundary> boundaries;
    std::vector<double> vertices;
    std::vector<uint32_t> triangles;
    std::vector<double> table;
    std::vector<owe_plane_wave> waves;
    owe_scene scene{};
};

// Throws std::runtime_error naming the first feature the hip backend does not implement (textures,
// stained glass, gradient or map skies, tabulated indices, cones, tori, round walls, domes, holed
// disks, diffuse aim guides, sensor focus guides).
void flattenForHip(const Scene& scene, const Detector& detector, HipFlatScene& out);

}  // namespace owe
