// Host-only driver of the file loaders (cella-nova_amd/csrc/loader.cpp) for a build under
// AddressSanitizer and UndefinedBehaviorSanitizer: every path on the command line is loaded as a
// creature config (.json, then written back and reloaded when it loads) or as a kernel file
// (.bin, at several kernel counts and target sizes), and the status of each call is printed.
// Any memory error or undefined behaviour aborts the process with the sanitizer's report.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "cellanova.h"

static double touch(const cn_config* c) {
  double acc = 0.0;
  const size_t plane = (size_t)c->seed_size * c->seed_size;
  for (size_t i = 0; i < (size_t)c->num_channels * plane; ++i) acc += c->seed_channels[i];
  for (uint32_t k = 0; k < c->num_kernels; ++k)
    acc += c->c0[k] + c->c1[k] + c->growth_mu[k] + c->growth_sigma[k] + c->growth_weights[k];
  return acc;
}

int main(int argc, char** argv) {
  double sink = 0.0;
  for (int i = 1; i < argc; ++i) {
    const std::string path = argv[i];
    if (path.size() > 5 && path.compare(path.size() - 5, 5, ".json") == 0) {
      cn_config* cfg = nullptr;
      const cn_status st = cn_load_config(path.c_str(), &cfg);
      int st2 = -1;
      if (st == CN_OK) {
        sink += touch(cfg);
        const std::string back = path + ".back.json";
        st2 = (int)cn_write_config(back.c_str(), cfg);
        cn_config* again = nullptr;
        if (st2 == CN_OK && cn_load_config(back.c_str(), &again) == CN_OK) {
          sink += touch(again);
          cn_free_config(again);
        }
        cn_free_config(cfg);
      }
      std::printf("%s %d %d\n", path.c_str(), (int)st, st2);
    } else {
      for (uint32_t K = 1; K <= 3; ++K)
        for (uint32_t target : {4u, 8u, 16u}) {
          std::vector<float> out(2 * (size_t)K * target * target);
          const cn_status st = cn_load_kernels(path.c_str(), K, target, out.data(), out.size());
          if (st == CN_OK)
            for (float v : out) sink += v;
          std::printf("%s K=%u target=%u %d\n", path.c_str(), K, target, (int)st);
        }
    }
  }
  std::printf("done %g\n", sink == sink ? 0.0 : 1.0);
  return 0;
}
