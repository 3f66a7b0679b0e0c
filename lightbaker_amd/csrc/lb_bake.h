// Host-side bake substrate (the reference's Rust host restated in C++).
#pragma once

#include <cstdint>
#include <vector>

#include "../../include/lightbaker.h"

namespace lbk {

// bvh.rs:50-106 over non-indexed triangles (9 floats each) with their material and chart.
void build_bvh(const float *positions, size_t n_tris, const uint32_t *material,
               const uint32_t *chart, const uint8_t *casts, std::vector<lb_node> &nodes,
               std::vector<lb_tri> &tris);

}  // namespace lbk
