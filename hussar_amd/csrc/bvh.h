// Host-side BVH2 builder for the gfx950 traversal kernels.
#pragma once

#include <stdint.h>
#include <vector>

namespace hsr {

// Flattened BVH ready for upload.
//  nodes: 16 floats per four-wide inner node (64 B) with 8-bit child boxes:
//         origin.xyz (float), biased exponents ex | ey << 8 | ez << 16 (uint32 bits) |
//         qlo.x, qhi.x, qlo.y, qhi.y (uint32 bits, byte i = child i) |
//         qlo.z, qhi.z, 0, 0 | ref[4] (int bits).
//         Child i's box is origin + q * 2^(e - 127) per axis, evaluated in float as a product
//         (exact) and a sum (rounded); the builder picks each q so that the decoded box
//         contains the child's widened box. Node 0 is the root. A ref >= 0 is an inner node,
//         kNoChild an unused slot, any other negative ref a leaf:
//         0x80000000 | (count-1) << 27 | firstSlot.
//  tris:  12 floats per triangle slot (48 B): v0.xyz, id (int bits) | e1.xyz, 0 | e2.xyz, 0
//         with e1 = v1 - v0 and e2 = v2 - v0 evaluated in float.
constexpr int32_t kNoChild = 0x7fffffff;
// traversal stack entries per ray the kernels provide (LDS part + private spill)
constexpr int kStackEntries = 16 + 128;

struct FlatBVH {
    std::vector<float> nodes;   // quantised four-wide nodes (closest-hit traversal)
    std::vector<float> nodesF;  // the same nodes with float child boxes, 32 floats each (any-hit)
    std::vector<float> tris;
    int32_t nodeCount = 0;
    int32_t slotCount = 0;
    int32_t maxDepth = 0;  // of the binary tree
    int32_t maxLeaf = 0;
    int32_t depth4 = 0;    // of the four-wide tree
};

// The four-wide collapse of a binary node array (16 floats per node, root 0) into 32-float
// nodes: lo.x[4] | hi.x[4] | lo.y[4] | hi.y[4] | lo.z[4] | hi.z[4] | ref[4] | unused.
void collapse4(const std::vector<float> &n2, std::vector<float> &n4, int32_t &depth4);
// 8-bit conservative quantisation of 32-float four-wide nodes into the 16-float layout above.
bool quantize4(const std::vector<float> &n4, std::vector<float> &out);
// origin + q * scale as the kernels evaluate it
float dequant(float origin, uint32_t q, float scale);

// Binned SAH (64 bins, all three axes) over triangle references (a triangle with a large box
// enters as several clipped references) with spatial splits where object splits overlap
// (bvh.cpp), leaves of up to 8 references, depth <= 80
// (median splits below depth 48). Child boxes are widened by a relative margin of 2^-20 plus
// 1e-6 m so that a ray accepted by the triangle test is never culled by a box test. Slots hold
// the referenced triangles in leaf order (slotCount >= triangleCount).
bool build_bvh(const float *vertices, int64_t vertexCount, const int32_t *indices, int64_t triangleCount,
               FlatBVH &out, int threads = 0);

}  // namespace hsr
