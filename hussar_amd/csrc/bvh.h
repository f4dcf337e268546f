// Host-side BVH2 builder for the gfx950 traversal kernels.
#pragma once

#include <stdint.h>
#include <vector>

namespace hsr {

// Flattened BVH ready for upload.
//  nodes: 32 floats per four-wide inner node (128 B): lo.x[4] | hi.x[4] | lo.y[4] | hi.y[4] |
//         lo.z[4] | hi.z[4] | ref[4] (int bits) | unused. Node 0 is the root. A ref >= 0 is an
//         inner node, kNoChild an unused slot, any other negative ref a leaf:
//         0x80000000 | (count-1) << 27 | firstSlot.
//  tris:  12 floats per triangle slot (48 B): v0.xyz, id (int bits) | e1.xyz, 0 | e2.xyz, 0
//         with e1 = v1 - v0 and e2 = v2 - v0 evaluated in float.
constexpr int32_t kNoChild = 0x7fffffff;
// traversal stack entries per ray the kernels provide (LDS part + private spill)
constexpr int kStackEntries = 16 + 128;

struct FlatBVH {
    std::vector<float> nodes;
    std::vector<float> tris;
    int32_t nodeCount = 0;
    int32_t slotCount = 0;
    int32_t maxDepth = 0;  // of the binary tree
    int32_t maxLeaf = 0;
    int32_t depth4 = 0;    // of the four-wide tree
};

// The four-wide collapse of a binary node array (16 floats per node, root 0).
void collapse4(const std::vector<float> &n2, std::vector<float> &n4, int32_t &depth4);

// Binned SAH (16 bins, all three axes), leaves of up to 8 triangles, depth <= 80 (median
// splits below depth 48). Child boxes are widened by a relative margin of 2^-20 plus 1e-6 m
// so that a ray accepted by the triangle test is never culled by a box test.
bool build_bvh(const float *vertices, int64_t vertexCount, const int32_t *indices, int64_t triangleCount,
               FlatBVH &out, int threads = 0);

}  // namespace hsr
