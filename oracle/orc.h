/*
 * Oracle: CPU restatement of the reference Piculator arithmetic.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (piculator_amd/) links,
 * loads or calls this code.  It is used by tests/ as the checker, by
 * __graft_entry__.smoke() as the checker, and by bench.py as the timed CPU
 * baseline ("port").
 *
 * Every function below restates one reference function; the citation next to
 * each declaration names the reference file:line it follows.  Reference
 * quirks that are undefined behaviour (NTT tables missing for k <= 26, no
 * primes for k < 11) are NOT inherited: tables exist for every k, and k < 11
 * uses the k = 11 prime (any prime with 2^k | p-1 gives the same product).
 */
#ifndef PICULATOR_ORACLE_H
#define PICULATOR_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Engine selection for the big multiply (BigFloat.cpp:379-441). */
enum orc_engine {
  ORC_ENGINE_REFERENCE = 0, /* FFT when k < 27, NTT otherwise (BigFloat.cpp:386) */
  ORC_ENGINE_NTT = 1,       /* every product through the NTT ("NTT-forced")      */
  ORC_ENGINE_FFT = 2        /* every product through the 8-bit FFT                */
};

/* Sign-magnitude float, BigFloat.h:33-36.  sign = 1 means positive or zero. */
typedef struct orc_bf {
  int32_t sign;
  int64_t exp;
  uint64_t L;
  uint32_t* T; /* malloc'd, L limbs little-endian, radix 2^32 */
} orc_bf;

/* ---- NTT engine (NTT.cpp) ---- */
uint64_t orc_ntt_prime(int k);                 /* primes[] table, NTT.cpp:125-158      */
uint64_t orc_ntt_generator(uint64_t p);        /* generator(), NTT.cpp:104-123         */
uint64_t orc_modmul(uint64_t a, uint64_t b, uint64_t p); /* NTT.cpp:85-90, 277-282    */
uint64_t orc_modpow(uint64_t a, uint64_t e, uint64_t p); /* NTT.cpp:91-100, 283-292   */
void orc_ntt_roots(int k, uint64_t* p, uint64_t* W, uint64_t* iW, uint64_t* rN); /* NNT::init, NTT.cpp:169-188 */
void orc_int_to_ntt(uint64_t* T, int k, const uint32_t* A, size_t AL);   /* NTT.cpp:330-341 */
void orc_ntt_forward(uint64_t* T, int k, int threads);                   /* NTT.cpp:306-313 */
void orc_ntt_inverse(uint64_t* T, int k, int threads);                   /* NTT.cpp:315-322 */
void orc_ntt_pointwise(uint64_t* T, const uint64_t* A, int k);           /* NTT.cpp:324-328 */
int orc_ntt_to_int(const uint64_t* T, int k, uint32_t* A, size_t AL);    /* NTT.cpp:343-369; -1 = "FFT length is too small." */

/* ---- FFT engine (FFT.cpp); complex values interleaved (re, im) ---- */
void orc_fft_ensure_table(int k);                                        /* FFT.cpp:31-60   */
void orc_int_to_fft(double* T, int k, const uint32_t* A, size_t AL);     /* FFT.cpp:485-519 */
void orc_fft_forward(double* T, int k, int threads);                     /* FFT.cpp:64-265  */
void orc_fft_inverse(double* T, int k, int threads);                     /* FFT.cpp:266-463 */
void orc_fft_pointwise(double* T, const double* A, int k);               /* FFT.cpp:464-484 */
int orc_fft_to_int(const double* T, int k, uint32_t* A, size_t AL);      /* FFT.cpp:520-571 */

/* Exact limb product through the chosen engine: out[0 .. na+nb) untrimmed.
 * Follows the body of BigFloat::mul, BigFloat.cpp:377-441 (sizing 380-385). */
int orc_mul_limbs(const uint32_t* a, size_t na, const uint32_t* b, size_t nb,
                  uint32_t* out, int engine, int threads);

/* ---- BigFloat (BigFloat.cpp) ---- */
void orc_bf_free(orc_bf* x);
void orc_bf_from_u32(orc_bf* z, uint32_t x, int sign);                    /* BigFloat.cpp:64-79   */
void orc_bf_copy(orc_bf* z, const orc_bf* x);                             /* BigFloat.cpp:24-33   */
int orc_bf_ucmp(const orc_bf* a, const orc_bf* b);                        /* BigFloat.cpp:123-147 */
void orc_bf_negate(orc_bf* a);                                            /* BigFloat.cpp:150-155 */
void orc_bf_mul_u32(orc_bf* z, const orc_bf* a, uint32_t x);              /* BigFloat.cpp:156-183 */
void orc_bf_add(orc_bf* z, const orc_bf* a, const orc_bf* b, size_t p);   /* BigFloat.cpp:288-305 */
void orc_bf_sub(orc_bf* z, const orc_bf* a, const orc_bf* b, size_t p);   /* BigFloat.cpp:306-325 */
void orc_bf_mul(orc_bf* z, const orc_bf* a, const orc_bf* b, size_t p, int engine, int threads); /* BigFloat.cpp:326-444 */
int orc_bf_rcp(orc_bf* z, const orc_bf* a, size_t p, int engine, int threads);                   /* BigFloat.cpp:445-511 */
int orc_bf_div(orc_bf* z, const orc_bf* a, const orc_bf* b, size_t p, int engine, int threads);  /* BigFloat.cpp:512-515 */
int orc_bf_invsqrt(orc_bf* z, uint32_t x, size_t p, int engine, int threads);                    /* BigFloat.cpp:516-567 */
/* Test self-checks only: 1 = p + 1 chop in the last sub of a Newton step, 2 = s = p / 2. */
void orc_set_plant(int k);
/* to_string_hex (BigFloat.cpp:95-106); returns bytes written (2 + digits) or -1. */
long orc_bf_to_hex(const orc_bf* a, size_t digits, char* out, size_t cap);

/* ---- Chudnovsky driver (PI.h) ---- */
/* Sizing of Pi(), PI.h:141-144. */
void orc_pi_sizes(size_t decimal_digits, size_t* hex_digits, size_t* p, size_t* terms);
/* Binary-splitting recursion, PI.h:64-112. */
void orc_pi_bsr(orc_bf* P, orc_bf* Q, orc_bf* R, uint32_t a, uint32_t b, size_t p,
                int engine, int threads);
/* Whole Pi(), PI.h:138-210: writes "3." + hex digits (the content of pi.txt).
 * times[5] (may be NULL) receives the phase wall times of PI.h:175-209:
 * series, division, invsqrt, final multiply, base conversion.
 * Returns the number of bytes written or -1. */
long orc_pi_hex(size_t decimal_digits, int engine, int threads, char* out, size_t cap,
                double* times);

/* Six hexadecimal digits of pi starting at digit n + 1 after the point, by
 * Bailey-Borwein-Plouffe digit extraction (Bailey, Borwein & Plouffe 1995):
 * frac(4 S1 - 2 S4 - S5 - S6) with S_j = sum_k 16^(n-k) / (8k + j).
 * Independent of the binary splitting; used to spot-check pi output.
 * Returns the digits as the top 24 bits of a 32-bit value. */
uint32_t orc_bbp_hex(uint64_t n, int threads);

/* Number of big (BigFloat x BigFloat) multiplies performed so far (statistics). */
uint64_t orc_mul_count(void);

#ifdef __cplusplus
}
#endif
#endif
