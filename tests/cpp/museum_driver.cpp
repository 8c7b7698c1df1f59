// museum_driver.cpp — compiles a scene script written for the reference (its
// scenes/museum_final.cpp, path in PRR_SCENE_SCRIPT) against the drop-in API
// through the forwarding headers of include/prr/reference, with the calls the
// reference's main.cpp makes, in its order (main.cpp:11-40). The script's
// assets are not shipped, so the driver is compiled, not run.
#include "Renderer.hpp"
#include "Scene.hpp"
#include "Sphere.hpp"
#include "ToneMapping.hpp"
#include "Triangle.hpp"
#include "Vector.hpp"
#include "global.hpp"
#include PRR_SCENE_SCRIPT

int main() {
    RenderSettings settings;
    settings.width = 1920 / 2;
    settings.height = 1080 / 2;
    settings.minSPP = 1;
    settings.maxSPP = 4096 / 16;
    settings.russianRoulette = 0.95f;
    settings.maxDepth = 12;
    settings.varianceThreshold = 0.01f;
    settings.exposure = 1.0f;
    settings.toneMapper = tonemap::ToneMapper::PBRNeutral;
    Scene scene = buildScene();
    scene.camera.init(settings.width, settings.height);
    scene.medium.sigma_t = -1;
    Integrator integrator = Integrator(scene, settings.maxDepth, settings.russianRoulette);
    scene.buildBVH();
    scene.buildPhotonMaps(1000000);
    Renderer r;
    r.Render(scene, integrator, settings);
    return 0;
}
