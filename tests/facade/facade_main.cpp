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
 *   facade_main <scene.json> <frames> <out.bin> op...
 *       first moves the camera as the visualizer's mouse handlers do
 *       (PinholeCamera::Rotate/Zoom/Pan, pinhole_camera.cpp:42-86), one op
 *       per argument: rotate,dphi,dtheta | zoom,d | pan,dx,dy; then prints
 *       the camera as hex floats ("camera phi theta radius tx ty tz up").
 */
#include "lantern/lantern.h"

#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>

int main(int argc, char **argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s <scene.json> <frames> <out.bin>\n", argv[0]);
        return 1;
    }
    Lantern::Scene scene;
    if (!scene.LoadSceneFromJSON(argv[1])) return 2;
    for (int k = 4; k < argc; ++k) {
        float a = 0.0f, b = 0.0f;
        if (std::sscanf(argv[k], "rotate,%f,%f", &a, &b) == 2) scene.Camera->Rotate(a, b);
        else if (std::sscanf(argv[k], "zoom,%f", &a) == 1) scene.Camera->Zoom(a);
        else if (std::sscanf(argv[k], "pan,%f,%f", &a, &b) == 2) scene.Camera->Pan(a, b);
        else {
            std::fprintf(stderr, "unknown camera op %s\n", argv[k]);
            return 1;
        }
    }
    if (argc > 4) {
        const lt_camera c = scene.Camera->Desc();
        std::printf("camera %a %a %a %a %a %a %a\n", c.phi, c.theta, c.radius, c.target[0], c.target[1], c.target[2],
                    c.up);
    }
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
