// ThreadSanitizer run of the oracle's thread pool (test infrastructure, like the oracle): the
// restatement compiled into this one translation unit with -fsanitize=thread (oracle/Makefile,
// target tsan_check), then driven through its C entry points the way the tests and the CPU
// baseline drive it: a guided schedule (training passes whose workers each fill their own
// statistics, a guide step between passes, a final pass), a MIMO render of two transmitters
// and four receivers with per-path hit and NEE records and counters, each on 8 workers
// sharing the batch counter of Backend::run (cpu.h:21-41). Prints the frames' sums; a data race
// makes ThreadSanitizer report and the process exit with its error code.
#include "hsr_oracle.cpp"

#include <cstdio>
#include <random>
#include <vector>

int main() {
    using namespace orc;
    // clutter: 600 random quads (1200 triangles) 4-12 m ahead of the radar, looking down +z
    std::mt19937 rng(7);
    std::uniform_real_distribution<float> u(0.0f, 1.0f);
    std::vector<float> V;
    std::vector<int32_t> I;
    for (int q = 0; q < 600; ++q) {
        const float cx = (u(rng) - 0.5f) * 6.0f, cy = (u(rng) - 0.5f) * 3.0f, cz = 4.0f + 8.0f * u(rng);
        const float ax[3] = {0.4f * u(rng) + 0.1f, 0.2f * (u(rng) - 0.5f), 0.2f * (u(rng) - 0.5f)};
        const float ay[3] = {0.2f * (u(rng) - 0.5f), 0.4f * u(rng) + 0.1f, 0.2f * (u(rng) - 0.5f)};
        const int32_t b = (int32_t)(V.size() / 3);
        const float p[4][3] = {{cx, cy, cz},
                               {cx + ax[0], cy + ax[1], cz + ax[2]},
                               {cx + ax[0] + ay[0], cy + ax[1] + ay[1], cz + ax[2] + ay[2]},
                               {cx + ay[0], cy + ay[1], cz + ay[2]}};
        for (auto &v : p) V.insert(V.end(), v, v + 3);
        const int32_t t[6] = {b, b + 1, b + 2, b, b + 2, b + 3};
        I.insert(I.end(), t, t + 6);
    }
    void *mesh = orc_mesh_create(V.data(), (int64_t)V.size() / 3, I.data(), (int64_t)I.size() / 3);

    SceneC sc{};
    sc.rf = Rf{77e9f, 60e12f, 10e6f, 5e-6f, 25.6e-6f, 0.0f};
    const float eye[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    auto antenna = [&](float x) {
        Antenna a{};
        a.pos[0] = x;
        for (int k = 0; k < 9; ++k) a.rot[k] = eye[k];
        return a;
    };
    sc.tx = antenna(0.0f);
    sc.rxCount = 4;
    for (int r = 0; r < 4; ++r) sc.rx[r] = antenna(0.002f * (float)(r + 1));
    Knobs kn{};
    kn.onlyIndirect = 1;
    kn.maxDepth = 3;
    kn.useGeometricalOptics = 1;
    kn.doGuiding = 1;
    kn.clearBeforeIteration = 1;
    kn.doFiltering = 1;
    kn.filteringSphere = 1;
    kn.filteringMin = 40.0f;
    kn.filteringMax = 300.0f;
    kn.filteringRadius = 160.0f;
    FrameCfg fc{1, 256, 4};
    std::vector<double> frame(2 * 256 * 4), frame2(2 * 256 * 8);
    double tw = 0.0;
    void *guide = orc_guide_create(0.2f, 0.01f, 1);
    int rc = orc_path_tracer_run_schedule(mesh, &sc, &kn, &fc, 1, 2, 4096, 8192, guide, 8, frame.data(), &tw);
    if (rc) return 10 + rc;

    // MIMO with records and counters, unguided
    sc.txCount = 2;
    sc.txMore[0] = antenna(-0.01f);
    kn.doGuiding = 0;
    FrameCfg fc2{1, 256, 8};
    const int64_t n = 4096;
    std::vector<int32_t> hits(n * 3);
    std::vector<int8_t> nee(n * 3 * 8);
    int64_t counters[5] = {};
    double tw2 = 0.0;
    rc = orc_render(mesh, &sc, &kn, &fc2, 1, 0, n, 1.0f, 1, nullptr, 8, frame2.data(), &tw2, hits.data(), nee.data(),
                    n, counters);
    if (rc) return 20 + rc;
    double s1 = 0.0, s2 = 0.0;
    for (double x : frame) s1 += x;
    for (double x : frame2) s2 += x;
    std::printf("schedule %.6e weight %.0f mimo %.6e weight %.0f nodes %lld\n", s1, tw, s2, tw2,
                (long long)counters[0]);
    orc_guide_destroy(guide);
    orc_mesh_destroy(mesh);
    return 0;
}
