// ref_cli_stubs.cpp -- the five renderer entry points the reference's main.cu declares
// (main.cu:8-12).  The renderer (render.cu: CUDA rasteriser + ffmpeg pipe) is not part of
// libplanes_hip.so, so a build of the reference's CLI against the library takes these in
// its place: each video mode reports that it is unavailable and exits 1, the way main.cu
// reports an unknown mode (main.cu:101-103).  Test infrastructure: linked only into the
// reference-CLI build of oracle/Makefile (target refcli).
#include <stdio.h>
#include <stdlib.h>

struct Genome;

static void no_renderer(const char* what) {
    fprintf(stderr, "%s: the renderer is not part of this build (libplanes_hip.so has no video path)\n", what);
    exit(1);
}

void render_showcase(const Genome*, int, const char*, int, const char*) { no_renderer("showcase"); }
void render_hero(const Genome*, const char*, int, const char*) { no_renderer("hero"); }
void render_robust(const Genome*, int, const char*, int, const char*) { no_renderer("robust-video"); }
void render_turntable(const Genome*, const char*, int, const char*) { no_renderer("turntable"); }
void render_evolution(const Genome*, int, const char*, int, const char*) { no_renderer("evolution-video"); }
