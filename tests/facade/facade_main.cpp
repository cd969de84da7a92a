/*
 * facade_main.cpp - drives the C++ facade (include/lantern/lantern.h) the way
 * Lantern's main.cpp drives its own classes (source/lantern/main.cpp:21-84):
 * Scene::LoadSceneFromJSON, FrameBuffer, Integrator::RenderFrame per frame.
 * Test infrastructure for tests/test_facade*.py.
 *
 *   facade_main <scene.json> <frames> <out.bin>
 *       renders `frames` frames and writes width, height (u32), then the
 *       FrameBuffer's ColorData (float3 per pixel), Bounces and
 *       ColorSampleCount (u32 per pixel). Exit 0, or 2 if the scene does not
 *       load (the loader's message is on stdout).
 */
#include "lantern/lantern.h"

#include <atomic>
#include <cstdio>
#include <cstdlib>

int main(int argc, char **argv) {
    if (argc != 4) {
        std::fprintf(stderr, "usage: %s <scene.json> <frames> <out.bin>\n", argv[0]);
        return 1;
    }
    Lantern::Scene scene;
    if (!scene.LoadSceneFromJSON(argv[1])) return 2;
    const unsigned frames = (unsigned)std::atoi(argv[2]);
    const unsigned w = scene.Camera->FrameBufferWidth, h = scene.Camera->FrameBufferHeight;
    /* RenderFrame swaps the integrator's buffer with *swap before rendering
     * (integrator.cpp:35); with *swap == the same buffer every frame lands in
     * it, as when the visualizer hands the accumulation buffer back */
    Lantern::FrameBuffer a(w, h);
    std::atomic<Lantern::FrameBuffer *> swap(&a);
    Lantern::Integrator integrator(&scene, &a, &swap);
    for (unsigned f = 0; f < frames; ++f) integrator.RenderFrame();
    Lantern::FrameBuffer *fb = &a;
    FILE *out = std::fopen(argv[3], "wb");
    if (!out) return 1;
    std::fwrite(&w, 4, 1, out);
    std::fwrite(&h, 4, 1, out);
    std::fwrite(fb->ColorData, sizeof(Lantern::float3), (size_t)w * h, out);
    std::fwrite(fb->Bounces, 4, (size_t)w * h, out);
    std::fwrite(fb->ColorSampleCount, 4, (size_t)w * h, out);
    std::fclose(out);
    std::printf("rendered %u frames of %ux%u\n", frames, w, h);
    return 0;
}
