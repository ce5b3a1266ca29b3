// Chudnovsky pi by binary splitting (reference PI.h:64-210) on the GPU
// BigFloat.  Two schedules produce byte-identical output:
//   Plain   -- the reference's depth-first recursion; every BigFloat::mul is
//              one GPU launch sequence (PI.h:64-112).
//   Batched -- the same tree (same midpoints, same truncation), evaluated
//              bottom-up by node height; all products of one height go to
//              the GPU as one batch (MulBatch).
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <string>

namespace piculator {

enum class PiMode { Plain = 0, Batched = 1 };

struct PiTimes {
  double series = 0, division = 0, invsqrt = 0, final_mul = 0, conversion = 0, total = 0;
  uint64_t big_muls = 0;
};

// Sizing of Pi() (PI.h:141-144).
void pi_sizes(size_t decimal_digits, size_t* hex_digits, size_t* p, size_t* terms);

// Content of pi.txt ("3." + hex digits) for Pi(decimal_digits).
std::string pi_hex(size_t decimal_digits, PiMode mode, PiTimes* times = nullptr,
                   bool verbose = false);

// The reference entry point (PI.h:138): prints the phase times and writes
// pi.txt in the working directory.
void Pi(size_t decimal_digits, size_t threads);

}  // namespace piculator
