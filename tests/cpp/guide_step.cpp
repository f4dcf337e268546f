// Host check of the product's guide step (hussar_amd/csrc/guide.cpp, GuidingWrapper::step of
// libhussar/include/hussar/integrators/path.h:119-126) on a given training tree.
// stdin:  splitThreshold secondMoment n
//         child[0..n-1]
//         stat[0..n-1]
// stdout: the sampling tree ("sampling m" then child[] and p0[] as %.9g), then the refined
//         training tree ("training m" then child[], depth[] and stat[]).
#include <cstdio>
#include <vector>

#include "../../hussar_amd/csrc/guide.h"

int main() {
    float thr = 0.0f;
    int sm = 1, n = 0;
    if (std::scanf("%f %d %d", &thr, &sm, &n) != 3 || n < 1) return 2;
    hsr::GuideHost g;
    g.splitThreshold = thr;
    g.secondMoment = sm != 0;
    hsr::GuideTree &t = g.training;
    t.child.assign(n, 0);
    t.p0.assign(n, 0.5f);
    t.stat.assign(n, 0.0);
    t.depth.assign(n, 0);
    for (int i = 0; i < n; ++i)
        if (std::scanf("%d", &t.child[i]) != 1) return 2;
    for (int i = 0; i < n; ++i)
        if (std::scanf("%lf", &t.stat[i]) != 1) return 2;
    for (int i = 0; i < n; ++i)
        if (t.child[i]) {
            if (t.child[i] + 1 >= n) return 2;
            t.depth[t.child[i]] = t.depth[i] + 1;
            t.depth[t.child[i] + 1] = t.depth[i] + 1;
        }
    g.step();
    const hsr::GuideTree &s = g.sampling;
    std::printf("sampling %d\n", s.size());
    for (int i = 0; i < s.size(); ++i) std::printf("%d ", s.child[i]);
    std::printf("\n");
    for (int i = 0; i < s.size(); ++i) std::printf("%.9g ", (double)s.p0[i]);
    std::printf("\ntraining %d\n", t.size());
    for (int i = 0; i < t.size(); ++i) std::printf("%d ", t.child[i]);
    std::printf("\n");
    for (int i = 0; i < t.size(); ++i) std::printf("%d ", t.depth[i]);
    std::printf("\n");
    for (int i = 0; i < t.size(); ++i) std::printf("%.17g ", t.stat[i]);
    std::printf("\n");
    return 0;
}
