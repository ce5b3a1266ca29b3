// Chudnovsky pi by binary splitting (reference PI.h:64-210) on the GPU
// BigFloat.  Two schedules produce byte-identical output:
//   Plain   -- the reference's depth-first recursion; every BigFloat::mul is
//              one GPU launch sequence (PI.h:64-112).
//   Batched -- the same tree (same midpoints, same truncation) with every
//              value resident in GPU memory: subtrees whose values fit the
//              precision window are evaluated level by level on the GPU
//              (one batched launch set per level, pcg_df_bsr_exact), the
//              nodes above them, the division and 1/sqrt(10005) run on
//              DeviceFloat, and the hex text is produced on the GPU.
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

// The same written into out[0..cap) (NUL-terminated); returns the length.
size_t pi_hex_into(size_t decimal_digits, PiMode mode, char* out, size_t cap,
                   PiTimes* times = nullptr, bool verbose = false);

// The reference entry point (PI.h:138): prints the phase times and writes
// pi.txt in the working directory.
void Pi(size_t decimal_digits, size_t threads);

}  // namespace piculator
