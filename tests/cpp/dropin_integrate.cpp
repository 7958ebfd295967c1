// The reference's per-sample render loop (examples/main.cpp:33-70) against
// the C++ drop-in: an OpenMP loop over pixels, a UniformSampler per pixel,
// jittered sensor coordinates, Camera::sampleRay, Integrator::integrate per
// sample, NaN / negative samples skipped, Image::addPixel and divide.  The
// same integrator then renders the frame in one device call (with the
// variance of each pixel mean) for the test to compare.  Arguments:
// model.obj width height spp {pt|pm}.  Prints the loop image, its variance,
// the device frame and its variance as hex floats, then a light sample.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "pm/render.hpp"

int main(int argc, char** argv) {
  if (argc != 6) {
    std::fprintf(stderr, "usage: %s model.obj width height spp pt|pm\n", argv[0]);
    return 2;
  }
  const int width = std::atoi(argv[2]), height = std::atoi(argv[3]);
  const int n_samples = std::atoi(argv[4]);
  const std::string mode = argv[5];
  try {
    Scene scene;
    scene.loadModel(argv[1]);
    scene.build();
    const float fov = 40.0f * PI / 180.0f;
    const float focal = 1.0f / std::tan(0.5f * fov);
    Camera camera(Vec3f(0.5f, 0.5f, 2.3f + focal), Vec3f(0.0f, 0.0f, -1.0f), fov);
    PathTracing pt(100);
    PhotonMapping pm(20000, 50, 1.0f, 50, 1, 100);
    Integrator& integrator = mode == "pt" ? static_cast<Integrator&>(pt) : pm;
    UniformSampler sampler;
    integrator.build(scene, sampler);

    Image image(width, height);
    Image sq(width, height);  // sums of squares, for the variance of each mean
    // The one change to the reference's loop: each pixel's sampler seed is
    // scrambled.  The reference seeds PCG32 with the pixel index, and PCG32's
    // first output is 0 for every state below 2^27, so sample 0 of every
    // pixel would sit on its left edge; the device frame it is compared with
    // jitters uniformly.
    auto scramble = [](uint64_t z) {
      z += 0x9e3779b97f4a7c15ull;
      z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
      z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
      return z ^ (z >> 31);
    };
#pragma omp parallel for collapse(2) schedule(dynamic, 1)
    for (int i = 0; i < height; ++i) {
      for (int j = 0; j < width; ++j) {
        UniformSampler sampler(scramble(j + width * i));
        for (int k = 0; k < n_samples; ++k) {
          const float u = (2.0f * (j + sampler.getNext1D()) - width) / height;
          const float v = (2.0f * (i + sampler.getNext1D()) - height) / height;
          Ray ray;
          float pdf;
          if (camera.sampleRay(Vec2f(u, v), ray, pdf)) {
            const Vec3f radiance = integrator.integrate(ray, scene, sampler) / pdf;
            if (std::isnan(radiance[0]) || std::isnan(radiance[1]) || std::isnan(radiance[2]))
              continue;
            if (radiance[0] < 0 || radiance[1] < 0 || radiance[2] < 0) continue;
            image.addPixel(i, j, radiance);
            sq.addPixel(i, j, radiance * radiance);
          } else {
            image.setPixel(i, j, Vec3f(0));
          }
        }
      }
    }
    image.divide(n_samples);
    sq.divide(n_samples);

    // the same integrator's device frame, with the variance of each mean
    std::vector<float> mean(3 * width * height), var(3 * width * height);
    if (mode == "pt") {
      pmx_render_params p{};
      p.width = width;
      p.height = height;
      p.spp = 4 * n_samples;
      p.max_depth = 100;
      p.seed = 77;
      p.out_mem = PM_MEM_HOST;
      pm::check(pmx_render_path_tracing(scene.handle(), &camera.desc(), &p, mean.data(),
                                        var.data(), nullptr),
                "pmx_render_path_tracing");
    } else {
      const pmx_render_params p = pm.frameParams(width, height, 4 * n_samples, 77);
      pm::check(pmx_render_photon_mapping(scene.handle(), &camera.desc(), pm.getPhotonMapPtr(),
                                          pm.getCausticsPhotonMapPtr(), &p, mean.data(),
                                          var.data(), nullptr),
                "pmx_render_photon_mapping");
    }
    const std::vector<float>& m = image.data();
    const std::vector<float>& q = sq.data();
    for (size_t c = 0; c < m.size(); ++c) std::printf("%a\n", m[c]);
    for (size_t c = 0; c < m.size(); ++c) {
      const double vv = (static_cast<double>(q[c]) - static_cast<double>(m[c]) * m[c]) /
                        (n_samples > 1 ? n_samples - 1 : 1);
      std::printf("%a\n", static_cast<float>(vv > 0 ? vv : 0.0));
    }
    for (float x : mean) std::printf("%a\n", x);
    for (float x : var) std::printf("%a\n", x);

    // Scene::sampleLight and the light's point and direction sampling
    UniformSampler ls(5);
    float pdf_choose, pdf_pos, pdf_dir;
    const auto light = scene.sampleLight(ls, pdf_choose);
    const SurfaceInfo sp = light->samplePoint(ls, pdf_pos);
    const Vec3f dir = light->sampleDirection(sp, ls, pdf_dir);
    const Vec3f le = light->Le(sp, dir);
    std::printf("%a %a %a %a %a %a %a %a %a %a %a %a\n", pdf_choose, pdf_pos, pdf_dir,
                sp.position[0], sp.position[1], sp.position[2], dir[0], dir[1], dir[2], le[0],
                le[1], le[2]);
  } catch (const pm::Error& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
