// Compile check: the front end's Backend and PathTracer accept types shaped like the
// reference's (Eigen-style vectors with operator[], matrices with operator()(r, c), a mesh of
// packed float3 / int3, a Scene with rfConfig / rx / tx). Also prints the vertex and index
// buffers of addQuad / addBox of the front end's own TriangleMesh, or, given file names, of
// the Wavefront OBJ files read one after another into one mesh, for a comparison.
#include "hussar_hip.hpp"

#include <cstdio>

namespace refshape {
struct Vec3 {
    float d[3];
    float operator[](int i) const { return d[i]; }
};
struct Mat3 {
    float d[9];
    float operator()(int r, int c) const { return d[3 * r + c]; }
};
struct Antenna {
    Vec3 p{{0, 0, 0}};
    Mat3 r{{1, 0, 0, 0, 1, 0, 0, 0, 1}};
    const Vec3 &position() const { return p; }
    const Mat3 &rotation() const { return r; }
};
struct RF {
    float startFreq, freqSlope, adcRate, idleTime, rampTime, antennaDelay;
};
struct Scene {
    RF rfConfig;
    Antenna rx, tx;
};
struct Mesh {
    struct Tri {
        int v0, v1, v2;
    };
    std::vector<Vec3> vertexBuffer;
    std::vector<Tri> indexBuffer;
};
}  // namespace refshape

// instantiated, never called: a run needs a device
void instantiate(const refshape::Mesh &mesh, const refshape::Scene &scene) {
    hussar::hip::PathTracer integrator;
    hussar::hip::Backend backend{mesh, integrator};
    integrator.run(backend, scene, 1024);
    bool stop = false;
    backend.run(scene, 16, &stop);
}

int main(int argc, char **argv) {
    hussar::hip::TriangleMesh m;
    if (argc > 1) {
        try {
            for (int a = 1; a < argc; ++a) hussar::hip::WavefrontFile(argv[a]).read(m);
        } catch (const hussar::hip::Error &e) {
            std::printf("error %s\n", e.what());
            return 1;
        }
        for (const auto &v : m.vertexBuffer) std::printf("v %a %a %a\n", v[0], v[1], v[2]);
        for (const auto &t : m.indexBuffer) std::printf("f %d %d %d\n", t.v0, t.v1, t.v2);
        return 0;
    }
    m.addQuad({0.5f, -1, 2}, {0, 0.25f, 0}, {0.125f, 0, 0});
    m.addBox({-1, 0, 0.5f}, {2, 0.75f, 3});
    for (const auto &v : m.vertexBuffer) std::printf("v %a %a %a\n", v[0], v[1], v[2]);
    for (const auto &t : m.indexBuffer) std::printf("f %d %d %d\n", t.v0, t.v1, t.v2);
    const hsr_scene s = hussar::hip::detail::scene(refshape::Scene{{1, 2, 3, 4, 5, 6}, {}, {}});
    std::printf("rf %g %g %g %g %g %g rx %d\n", s.rfConfig.startFreq, s.rfConfig.freqSlope, s.rfConfig.adcRate,
                s.rfConfig.idleTime, s.rfConfig.rampTime, s.rfConfig.antennaDelay, s.rxCount);
    return 0;
}
