// C entry points of the host library libpiculator.so (for Python tests and
// bench.py): pi computation and single BigFloat operations on the GPU
// BigFloat.  C++ exceptions (the reference's string literals, and any
// std::exception such as std::bad_alloc from a host allocation) are turned
// into a status code and kept for pch_last_error(); none crosses extern "C".
#include <cstring>
#include <exception>
#include <new>
#include <string>

#include "../../include/picugpu.h"
#include "BigFloat.h"
#include "DeviceFloat.h"
#include "Pi.h"

using piculator::BigFloat;

namespace {
thread_local std::string g_err;

int catch_all(const char* what) {
  g_err = what ? what : "unknown error";
  return -1;
}

int export_bf(const BigFloat& z, int* sign, int64_t* exp, uint32_t* limbs, uint64_t cap,
              uint64_t* n) {
  *n = z.get_precision();
  *sign = z.get_sign() ? 1 : 0;
  *exp = z.get_exponent();
  if (*n > cap) return catch_all("output capacity too small");
  if (*n) std::memcpy(limbs, z.limbs(), *n * sizeof(uint32_t));
  return 0;
}
}  // namespace

extern "C" {

const char* pch_last_error(void) { return g_err.c_str(); }

// Selects the engine flags used by BigFloat::mul (PCG_ENGINE_*).
void pch_set_engine(uint32_t flags) { piculator::set_mul_engine_flags(flags); }

// "3." + hex digits of Pi(digits) (PI.h:138-210); mode 0 plain, 1 batched.
// times[7]: series, division, invsqrt, final multiply, conversion, total, big multiplies.
// Hex digits, precision in limbs and series terms of Pi(digits) (PI.h:141-144).
void pch_pi_sizes(uint64_t digits, uint64_t* hex_digits, uint64_t* p, uint64_t* terms) {
  size_t h, pp, t;
  piculator::pi_sizes(digits, &h, &pp, &t);
  *hex_digits = h;
  *p = pp;
  *terms = t;
}

long pch_pi_hex(uint64_t digits, int mode, char* out, uint64_t cap, double* times, int verbose) {
  try {
    piculator::PiTimes t;
    const size_t n = piculator::pi_hex_into(
        digits, mode == 0 ? piculator::PiMode::Plain : piculator::PiMode::Batched, out, cap, &t,
        verbose != 0);
    if (times) {
      times[0] = t.series;
      times[1] = t.division;
      times[2] = t.invsqrt;
      times[3] = t.final_mul;
      times[4] = t.conversion;
      times[5] = t.total;
      times[6] = (double)t.big_muls;
    }
    return (long)n;
  } catch (const char* e) {
    return catch_all(e);
  } catch (const std::bad_alloc&) {
    return catch_all("out of host memory");
  } catch (const std::exception& e) {
    return catch_all(e.what());
  } catch (...) {
    return catch_all("unknown C++ exception");
  }
}

// One BigFloat operation: op 0 mul(b, p), 1 add(b, p), 2 sub(b, p), 3 rcp(p),
// 4 div(b, p), 5 invsqrt(a_limbs[0], p), 6 mul(uint32 b_limbs[0]).
int pch_bf_op(int op, int sa, int64_t ea, const uint32_t* la, uint64_t na, int sb, int64_t eb,
              const uint32_t* lb, uint64_t nb, uint64_t p, int* sz, int64_t* ez, uint32_t* lz,
              uint64_t cap, uint64_t* nz) {
  try {
    const BigFloat a = BigFloat::from_parts(sa != 0, ea, la, na);
    const BigFloat b = BigFloat::from_parts(sb != 0, eb, lb, nb);
    BigFloat z;
    switch (op) {
      case 0: z = a.mul(b, p); break;
      case 1: z = a.add(b, p); break;
      case 2: z = a.sub(b, p); break;
      case 3: z = a.rcp(p); break;
      case 4: z = a.div(b, p); break;
      case 5: z = piculator::invsqrt(na ? la[0] : 0, p); break;
      case 6: z = a.mul(nb ? lb[0] : 0u); break;
      default: return catch_all("bad op");
    }
    return export_bf(z, sz, ez, lz, cap, nz);
  } catch (const char* e) {
    return catch_all(e);
  } catch (const std::bad_alloc&) {
    return catch_all("out of host memory");
  } catch (const std::exception& e) {
    return catch_all(e.what());
  } catch (...) {
    return catch_all("unknown C++ exception");
  }
}

// The same operations on DeviceFloat (values resident in device memory;
// uploaded, computed on the GPU, downloaded).  op 7: ucmp(a, b) into *sz.
int pch_df_op(int op, int sa, int64_t ea, const uint32_t* la, uint64_t na, int sb, int64_t eb,
              const uint32_t* lb, uint64_t nb, uint64_t p, int* sz, int64_t* ez, uint32_t* lz,
              uint64_t cap, uint64_t* nz) {
  try {
    using piculator::DeviceFloat;
    const DeviceFloat a(BigFloat::from_parts(sa != 0, ea, la, na));
    const DeviceFloat b(BigFloat::from_parts(sb != 0, eb, lb, nb));
    DeviceFloat z;
    switch (op) {
      case 0: z = a.mul(b, p); break;
      case 1: z = a.add(b, p); break;
      case 2: z = a.sub(b, p); break;
      case 3: z = a.rcp(p); break;
      case 4: z = a.div(b, p); break;
      case 5: z = piculator::invsqrt_device(na ? la[0] : 0, p); break;
      case 6: z = a.mul(nb ? lb[0] : 0u); break;
      case 7: {
        int cmp = 0;
        if (pcg_df_ucmp(&a.raw(), &b.raw(), &cmp) != PCG_OK) return catch_all(pcg_last_error_string());
        *sz = cmp;
        *ez = 0;
        *nz = 0;
        return 0;
      }
      default: return catch_all("bad op");
    }
    return export_bf(z.to_host(), sz, ez, lz, cap, nz);
  } catch (const char* e) {
    return catch_all(e);
  } catch (const std::bad_alloc&) {
    return catch_all("out of host memory");
  } catch (const std::exception& e) {
    return catch_all(e.what());
  } catch (...) {
    return catch_all("unknown C++ exception");
  }
}

}  // extern "C"
