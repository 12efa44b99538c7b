// Compiles against the C++ facade (include/cvo/CvoGPU.hpp) and links the
// library.  Without arguments it exercises the host-only parts; with a YAML
// path, --gpu and an output file it also registers two small clouds on the GPU
// and writes the clouds and the transform there.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "cvo/CvoGPU.hpp"

int main(int argc, char** argv) {
  std::printf("library %s\n", cvo_version());
  cvo_params p;
  cvo_params_default(&p);
  if (argc > 1) {
    cvo_status s = cvo_params_read_yaml(argv[1], &p);
    if (s != CVO_OK) {
      std::printf("yaml: %s\n", cvo_status_string(s));
      return 1;
    }
    std::printf("ell_init %g sp_thres %g MAX_ITER %d\n", p.ell_init, p.sp_thres, p.MAX_ITER);
  }
  // a plane patch and the same patch shifted by 2 cm
  cvo::CvoPointCloud a(5, 0), b(5, 0);
  const int n = 400;
  a.reserve(n, 5, 0);
  b.reserve(n, 5, 0);
  const float geo[2] = {0.0f, 1.0f};
  for (int i = 0; i < n; ++i) {
    const float x = 0.05f * (float)(i % 20), y = 0.05f * (float)(i / 20);
    const float pa[3] = {x, y, 0.1f * std::sin(3.0f * x)}, pb[3] = {x + 0.02f, y, 0.1f * std::sin(3.0f * x)};
    const float f[5] = {0.5f + 0.5f * std::sin(x), 0.5f, 0.5f, 0.5f, 0.5f};
    if (a.add_point(i, pa, f, nullptr, geo) != 0 || b.add_point(i, pb, f, nullptr, geo) != 0) return 2;
  }
  const cvo_cloud_view v = a.view();
  if (v.n != n || v.feat_dim != 5 || v.sem != nullptr || v.geo_type != nullptr) return 3;
  if (a.add_point(n, a.xyz_at(0), nullptr, nullptr, geo) != -1) return 4;
  if (argc > 3 && std::string(argv[2]) == "--gpu") {
    cvo::CvoGPU cvo(argv[1]);
    const float I[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    float T[16];
    double secs = 0.0;
    int iters = 0;
    const int ret = cvo.align(a, b, I, T, nullptr, &secs, &iters);
    const float c = cvo.function_angle(a, b, I, 0.1f);
    std::printf("align ret %d iters %d seconds %.4f cos %.6f\n", ret, iters, secs, c);
    // the clouds and the result, for a comparison with the other bindings
    FILE* f = std::fopen(argv[3], "wb");
    if (!f) return 5;
    const int32_t hdr[3] = {n, ret, iters};
    std::fwrite(hdr, sizeof(hdr), 1, f);
    for (const cvo::CvoPointCloud* c : {&a, &b}) {
      const cvo_cloud_view w = c->view();
      std::fwrite(w.xyz, sizeof(float), 3 * (size_t)n, f);
      std::fwrite(w.feat, sizeof(float), 5 * (size_t)n, f);
    }
    std::fwrite(T, sizeof(float), 16, f);
    std::fclose(f);
  }
  std::printf("ok\n");
  return 0;
}
