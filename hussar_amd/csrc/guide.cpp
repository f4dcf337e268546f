// Host build/refine of the guiding tree (GuidingWrapper::step, path.h:119-126).
#include "guide.h"

#include <algorithm>
#include <cmath>

namespace hsr {

static constexpr int32_t kGuideMaxDepth = 24;
static constexpr int32_t kGuideMaxNodes = 1 << 16;

void GuideTree::reset() {
    child.assign(1, 0);
    p0.assign(1, 0.5f);
    stat.assign(1, 0.0);
    depth.assign(1, 0);
}

void GuideHost::reset() {
    sampling.reset();
    training.reset();
}

void GuideHost::step() {
    GuideTree &t = training;
    const int32_t n = t.size();
    std::vector<double> mass(n, 0.0);
    // children follow their parent in index order, so one backward sweep sees children first
    for (int32_t i = n - 1; i >= 0; --i) {
        if (t.child[i] == 0) {
            const double s = std::max(t.stat[i], 0.0);
            mass[i] = secondMoment ? std::sqrt(s * std::ldexp(1.0, -t.depth[i])) : s;
        } else {
            const double m0 = mass[t.child[i]], m1 = mass[t.child[i] + 1];
            mass[i] = m0 + m1;
            t.p0[i] = mass[i] > 0.0 ? (float)(m0 / mass[i]) : 0.5f;
        }
    }
    sampling = t;
    const double total = mass[0];
    if (total > 0.0) {
        // shares of the new nodes: half of the parent's (statistics spread evenly over a leaf)
        std::vector<double> share(mass.size());
        for (size_t i = 0; i < mass.size(); ++i) share[i] = mass[i] / total;
        for (int32_t i = 0; i < t.size(); ++i) {
            if (t.child[i] != 0 || t.depth[i] >= kGuideMaxDepth) continue;
            if (t.size() + 2 > kGuideMaxNodes) break;
            if (share[i] > (double)splitThreshold) {
                const int32_t c = t.size();
                t.child[i] = c;
                t.p0[i] = 0.5f;
                for (int q = 0; q < 2; ++q) {
                    t.child.push_back(0);
                    t.p0.push_back(0.5f);
                    t.stat.push_back(0.0);
                    t.depth.push_back(t.depth[i] + 1);
                    share.push_back(share[i] * 0.5);
                }
            }
        }
    }
    std::fill(t.stat.begin(), t.stat.end(), 0.0);
}

}  // namespace hsr
