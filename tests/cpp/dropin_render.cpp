// Drives the C++ drop-in renderer (include/pm/render.hpp) the way the
// reference's examples/main.cpp drives its own: load an OBJ scene, build it,
// build the integrator, render.  Arguments: model.obj width height spp
// {pt|pm} seed.  Prints the camera record and the image as hex floats, then
// (pm) the numbers of stored photons, for the test to compare with the same
// render through the Python binding.
#include <cstdio>
#include <cstdlib>
#include <string>

#include "pm/render.hpp"

int main(int argc, char** argv) {
  if (argc != 7) {
    std::fprintf(stderr, "usage: %s model.obj width height spp pt|pm seed\n", argv[0]);
    return 2;
  }
  const int width = std::atoi(argv[2]), height = std::atoi(argv[3]), spp = std::atoi(argv[4]);
  const std::string mode = argv[5];
  const uint64_t seed = std::strtoull(argv[6], nullptr, 10);
  try {
    Scene scene;
    scene.loadModel(argv[1]);
    scene.build();
    // the box scene's camera: pinhole at (0.5, 0.5, 2.3), 40 degree vertical FOV
    const float fov = 40.0f * pm::kPi / 180.0f;
    const float focal = 1.0f / std::tan(0.5f * fov);
    Camera camera(pm::Vec3f(0.5f, 0.5f, 2.3f + focal), pm::Vec3f(0.0f, 0.0f, -1.0f), fov);
    const pmx_camera& c = camera.desc();
    for (const float* f : {c.position, c.forward, c.right, c.up})
      std::printf("%a %a %a ", f[0], f[1], f[2]);
    std::printf("%a\n", c.focal);
    // a few closest hits through Scene::intersect, printed as "prim t"
    for (int i = 0; i < 4; ++i) {
      float uv[2] = {-0.75f + 0.5f * i, 0.1f * i}, pdf;
      Ray ray;
      camera.sampleRay(uv, ray, pdf);
      IntersectInfo info;
      if (scene.intersect(ray, info)) std::printf("%d %a ", info.primID, info.t);
      else std::printf("-1 0 ");
    }
    std::printf("\n");
    UniformSampler sampler(seed);
    if (mode == "pt") {
      PathTracing integrator(100);
      integrator.build(scene, sampler);
      const Image img = integrator.render(scene, camera, width, height, spp, seed);
      for (float x : img.data()) std::printf("%a\n", x);
    } else {
      PhotonMapping integrator(20000, 50, 1.0f, 50, 1, 100);
      integrator.build(scene, sampler);
      Image img = integrator.render(scene, camera, width, height, spp, seed);
      for (float x : img.data()) std::printf("%a\n", x);
      std::printf("%lld %lld\n", static_cast<long long>(integrator.nGlobalPhotonsStored()),
                  static_cast<long long>(integrator.nCausticsPhotonsStored()));
      img.gammaCorrection(2.2f);
      img.writePPM("dropin_render.ppm");
    }
  } catch (const pm::Error& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
