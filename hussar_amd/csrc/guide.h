// Host side of the path-guiding distribution over primary sample space.
//
// The reference wraps libguiding's BTree<2, Leaf<Empty>> in GuidingWrapper
// (libhussar/include/hussar/integrators/path.h:17-131); libguiding's source is not part of the
// reference, so this is the backend's own 2-D binary tree with the wrapper's semantics:
//  * node 0 covers [0,1)^2; an inner node halves its cell along axis (depth % 2); the two
//    children are stored consecutively (child, child + 1), child 0 being the lower half;
//  * the sampling tree stores per inner node p0 = probability of child 0;
//  * the training tree accumulates per leaf Σ density^k · weight (k = 2 with secondMoment);
//  * step(): leaf mass m = sqrt(stat · area) (secondMoment) or stat; inner mass = sum of the
//    children; p0 = m(child 0) / m(node) (0.5 when the node carries no mass); the sampling
//    tree becomes this tree; then, in node order including the nodes this pass creates, every
//    leaf whose mass share exceeds splitThreshold (depth below 24, fewer than 65536 nodes) is
//    split in two, each child taking half of its parent's share (the SD-tree rebuild of
//    Müller et al. 2017 for a binary tree), and all statistics are cleared.
// Lookup (warp + pdf) and record (leaf statistics) run on the device.
#pragma once

#include <stdint.h>
#include <vector>

namespace hsr {

struct GuideTree {
    std::vector<int32_t> child{0};
    std::vector<float> p0{0.5f};
    std::vector<double> stat{0.0};
    std::vector<int32_t> depth{0};
    void reset();
    int32_t size() const { return (int32_t)child.size(); }
};

struct GuideHost {
    float uniformProb = 0.1f;
    float splitThreshold = 0.005f;
    bool secondMoment = true;
    GuideTree sampling;
    GuideTree training;
    void reset();
    void step();
};

}  // namespace hsr
