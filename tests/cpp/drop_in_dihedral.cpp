// The dihedral sweep of examples/src/1_dihedral.cpp:40-130 written against the header-only
// C++ front end: the reference's mesh / scene / integrator / backend calls with the backend
// type replaced. Two 10 cm plates meet along +y at the origin; a co-located Tx/Rx at 1.5 m
// on the bisector is yawed about +y. Every frame is appended to a .SIM file
// (examples/src/utils.h:29-40), and each pose's antenna values are printed as hex floats so
// that a test can rebuild the identical scene.
//
// usage: drop_in_dihedral <out.SIM> <samples> <doGuiding 0|1> <yaw_deg>...
#include "hussar_hip.hpp"

#include <cstdio>
#include <cstdlib>
#include <fstream>

using namespace hussar::hip;

int main(int argc, char **argv) {
    if (argc < 5) {
        std::fprintf(stderr, "usage: %s out.SIM samples doGuiding yaw_deg...\n", argv[0]);
        return 2;
    }
    const long samples = std::atol(argv[2]);

    RFConfig rf;
    rf.startFreq = 77e9f;
    rf.freqSlope = 60e12f;
    rf.adcRate = 10e6f;
    rf.idleTime = 100e-6f;
    rf.rampTime = 25.6e-6f;
    rf.antennaDelay = 0.0f;

    FrameConfig frameConfig;
    frameConfig.chirpCount = 1;
    frameConfig.samplesPerChirp = 256;
    frameConfig.channelCount = 1;

    TriangleMesh mesh;
    const float size = 0.1f;
    mesh.addQuad({0, 0, 0}, {0, size, 0}, {0, 0, size});
    mesh.addQuad({0, 0, 0}, {size, 0, 0}, {0, size, 0});

    Scene scene;
    scene.rfConfig = rf;

    PathTracer integrator;
    integrator.maxDepth = 3;
    integrator.doGuiding = std::atoi(argv[3]) != 0;
    integrator.configureFrame(frameConfig);

    try {
        Backend backend{mesh, integrator};
        std::ofstream file(argv[1], std::ios::binary);
        const Matrix33f facing = Matrix33f::rows(0, 0, -1, 0, -1, 0, -1, 0, 0);
        for (int a = 4; a < argc; ++a) {
            const float yaw = std::strtof(argv[a], nullptr);
            const Matrix33f rotation = Matrix33f::rotationY((yaw - 45.0f) / 180.0f * 3.14159265358979f);
            const NFAntenna antenna{rotation * Vector3f(1.5f, 0.05f, 0.0f), rotation * facing};
            scene.rx = antenna;
            scene.tx = antenna;
            integrator.run(backend, scene, samples);
            const Frame frame = integrator.fetchFrame();
            writeFrameToFile(file, frame);
            std::printf("pose %g peak %zu pos %a %a %a rot", yaw, frame.argmax(), antenna.position()[0],
                        antenna.position()[1], antenna.position()[2]);
            for (int i = 0; i < 9; ++i) std::printf(" %a", antenna.rotation().m[i]);
            std::printf("\n");
        }
    } catch (const Error &e) {
        std::fprintf(stderr, "error %d: %s\n", (int)e.status, e.what());
        return 1;
    }
    return 0;
}
