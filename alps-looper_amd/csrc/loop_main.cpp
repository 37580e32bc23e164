// loopgpu_run: run one "loop; path integral" task on the GPU from a parameter
// file in the reference's .ip syntax (or stdin) and print the "[results]"
// block -- the analogue of the reference's `loop` program for this path.
//
//   loopgpu_run params.ip
//   echo 'LATTICE="chain lattice"; L=16; T=0.1; SWEEPS=65536' | loopgpu_run
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/loopgpu_worker.h"

int main(int argc, char** argv) {
  std::stringstream ss;
  if (argc > 1) {
    std::ifstream f(argv[1]);
    if (!f) { std::cerr << "cannot open " << argv[1] << "\n"; return 1; }
    ss << f.rdbuf();
  } else {
    ss << std::cin.rdbuf();
  }
  std::cout << "[input parameters]\n" << ss.str() << "\n";
  lg_worker* w = nullptr;
  if (lg_worker_create(ss.str().c_str(), &w) != LG_OK) {
    char buf[1024];
    lg_worker_last_error(buf, sizeof buf);
    std::cerr << "error: " << buf << "\n";
    return 2;
  }
  int can = 1;
  while (can) {
    uint64_t done = 0;
    if (lg_worker_run(w, 8192, &done) != LG_OK) {
      char buf[1024];
      lg_worker_last_error(buf, sizeof buf);
      std::cerr << "error: " << buf << "\n";
      return 3;
    }
    uint64_t mcs;
    double progress;
    lg_worker_status(w, &mcs, nullptr, &progress, &can);
    std::cerr << "progress " << (int)(100 * progress) << "% (" << mcs << " sweeps)\n";
  }
  size_t need = 0;
  lg_worker_report(w, nullptr, 0, &need);
  std::vector<char> out(need);
  lg_worker_report(w, out.data(), out.size(), nullptr);
  std::cout << out.data();
  lg_worker_destroy(w);
  return 0;
}
