// Command-line front end (the reference's menu, main.cpp:8-41, as flags):
//   piculator [--digits D] [--mode batched|plain] [--out pi.txt]
// Computes pi to D decimal digits (hexadecimal output, as the reference) on
// the GPU and writes pi.txt.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/picugpu.h"
#include "Pi.h"

int main(int argc, char** argv) {
  size_t digits = 10000000;  // main.cpp:5
  std::string mode = "batched", out = "pi.txt";
  for (int i = 1; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--digits") && i + 1 < argc)
      digits = std::strtoull(argv[++i], nullptr, 10);
    else if (!std::strcmp(argv[i], "--mode") && i + 1 < argc)
      mode = argv[++i];
    else if (!std::strcmp(argv[i], "--out") && i + 1 < argc)
      out = argv[++i];
    else {
      std::fprintf(stderr, "usage: %s [--digits D] [--mode batched|plain] [--out FILE]\n", argv[0]);
      return 2;
    }
  }
  if (pcg_init(0) != PCG_OK) {
    std::fprintf(stderr, "GPU init failed: %s\n", pcg_last_error_string());
    return 1;
  }
  char* buf = nullptr;  // pi.txt in page-locked memory (the text comes off the GPU)
  try {
    size_t hex = 0, p = 0, terms = 0;
    piculator::pi_sizes(digits, &hex, &p, &terms);
    const size_t cap = hex + 32;
    if (pcg_host_alloc(reinterpret_cast<void**>(&buf), cap) != PCG_OK) throw pcg_last_error_string();
    piculator::PiTimes t;
    const size_t n = piculator::pi_hex_into(
        digits, mode == "plain" ? piculator::PiMode::Plain : piculator::PiMode::Batched, buf, cap, &t,
        true);
    std::printf("Piculator (MI355X)\nDecimal Digits: %zu  Hex Digits: %zu  mode: %s\n", digits, n - 2,
                mode.c_str());
    std::printf("series %.3f s, division %.3f s, invsqrt %.3f s, final multiply %.3f s, "
                "conversion %.3f s, total %.3f s, big multiplies %llu\n",
                t.series, t.division, t.invsqrt, t.final_mul, t.conversion, t.total,
                (unsigned long long)t.big_muls);
    FILE* f = std::fopen(out.c_str(), "wb");
    if (!f) throw "Cannot Create File";
    std::fwrite(buf, 1, n, f);
    std::fclose(f);
  } catch (const char* e) {
    std::fprintf(stderr, "error: %s\n", e);
    if (buf) pcg_host_free(buf);
    return 1;
  }
  pcg_host_free(buf);
  pcg_shutdown();
  return 0;
}
