// Host-side BVH2 builder for the gfx950 traversal kernels.
#pragma once

#include <stdint.h>
#include <vector>

namespace hsr {

// Flattened BVH ready for upload.
//  nodes: 16 floats per four-wide inner node (64 B) with 8-bit child boxes:
//         origin.xyz, scale.x (float) |
//         qlo.x, qhi.x, qlo.y, qhi.y (uint32 bits, byte i = child i) |
//         qlo.z, qhi.z (uint32 bits), scale.y, scale.z (float) | ref[4] (int bits).
//         Child i's box is origin + q * scale per axis, scale a power of two in [2^-120, 2^120],
//         evaluated in float as a product (exact) and a sum (rounded); the builder picks each q
//         so that the decoded box contains the child's widened box. Node 0 is the root. A
//         ref >= 0 is an inner node, any negative ref a leaf:
//         0x80000000 | (count-1) << 27 | firstSlot. An unused child slot has the inverted box
//         qlo = 255, qhi = 0 on every axis, which the box test rejects whenever the node's
//         scale is not tiny against its distance from the ray origin, and the sentinel leaf
//         (one slot at slotCount), so a box test that lets it through tests one degenerate
//         triangle, which nothing hits.
//  tris:  12 floats per triangle slot (48 B): v0.xyz, id (int bits) | e1.xyz, 0 | e2.xyz, 0
//         with e1 = v1 - v0 and e2 = v2 - v0 evaluated in float; slotCount slots, then the
//         sentinel (all zero, id -1).
constexpr int32_t kNoChild = 0x7fffffff;  // an unused child slot of the float nodes
// traversal stack entries per ray the kernels provide (LDS part + private spill)
constexpr int kStackEntries = 16 + 128;

struct FlatBVH {
    std::vector<float> nodes;   // quantised four-wide nodes (closest-hit traversal)
    std::vector<float> nodesF;  // the same nodes with float child boxes, 32 floats each (any-hit)
    std::vector<float> tris;    // slotCount + 1 slots (the last one the sentinel)
    int32_t nodeCount = 0;
    int32_t slotCount = 0;
    int32_t maxDepth = 0;  // of the binary tree
    int32_t maxLeaf = 0;
    int32_t depth4 = 0;    // of the four-wide tree
};

// The four-wide collapse of a binary node array (16 floats per node, root 0) into 32-float
// nodes: lo.x[4] | hi.x[4] | lo.y[4] | hi.y[4] | lo.z[4] | hi.z[4] | ref[4] | unused.
void collapse4(const std::vector<float> &n2, std::vector<float> &n4, int32_t &depth4);
// 8-bit conservative quantisation of 32-float four-wide nodes into the 16-float layout above;
// unused child slots become the inverted box with child reference emptyRef.
bool quantize4(const std::vector<float> &n4, std::vector<float> &out, int32_t emptyRef);
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
