// The C-ABI of the host substrate (include/lightbaker.h, second half).
#include <new>
#include <vector>

#include "../../include/lightbaker.h"
#include "lb_bake.h"

namespace lbk {
int fail(const char *fmt, ...);
}

struct lb_bvh {
    std::vector<lb_node> nodes;
    std::vector<lb_tri> tris;
};

extern "C" {

int lb_bvh_build(const float *positions, size_t n_tris, const uint32_t *material,
                 const uint32_t *chart, const uint8_t *casts, size_t n_materials, lb_bvh **out) {
    if (!out) return lbk::fail("lb_bvh_build: null handle");
    *out = nullptr;
    if (n_tris && (!positions || !material || !chart)) return lbk::fail("lb_bvh_build: null input");
    if (!casts && n_materials) return lbk::fail("lb_bvh_build: null casts");
    for (size_t t = 0; t < n_tris; t++)
        if (material[t] >= n_materials)
            return lbk::fail("lb_bvh_build: triangle %zu: material %u of %zu", t, material[t],
                             n_materials);
    auto *b = new (std::nothrow) lb_bvh();
    if (!b) return lbk::fail("lb_bvh_build: out of memory");
    lbk::build_bvh(positions, n_tris, material, chart, casts, b->nodes, b->tris);
    *out = b;
    return 0;
}

size_t lb_bvh_nodes(const lb_bvh *bvh, const lb_node **nodes) {
    if (!bvh) return 0;
    if (nodes) *nodes = bvh->nodes.data();
    return bvh->nodes.size();
}

size_t lb_bvh_tris(const lb_bvh *bvh, const lb_tri **tris) {
    if (!bvh) return 0;
    if (tris) *tris = bvh->tris.data();
    return bvh->tris.size();
}

void lb_bvh_destroy(lb_bvh *bvh) { delete bvh; }

}  // extern "C"
