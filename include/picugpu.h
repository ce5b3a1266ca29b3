/*
 * picugpu: MI355X (gfx950) big-integer multiply engine behind a plain C ABI.
 *
 * This library replaces the limb product inside BigFloat::mul
 * (reference BigFloat.cpp:377-441) -- the FFT/NTT convolution engines of
 * FFT.h:4-9 / NTT.h:4-9 and their calls -- with HIP kernels.  The metadata
 * logic of BigFloat::mul (zero test, precision chop, sign, exponent and the
 * one-limb trim, BigFloat.cpp:333-375,411-412,436-437) stays with the caller;
 * piculator_amd/host/BigFloat.cpp is such a caller.
 *
 * Limbs are little-endian uint32_t in radix 2^32, as BigFloat::T
 * (BigFloat.h:36).  Every entry point computes the EXACT integer product of
 * the two limb vectors; the result never depends on the engine, the primes,
 * the transform length or batching.
 *
 * Conventions: no C++ types cross this boundary; every function returns a
 * pcg_status; pcg_last_error_string() describes the last failure of the
 * calling thread.  Calls are thread-safe: they serialise on one internal
 * lock per library instance (one GPU, one stream).
 */
#ifndef PICUGPU_H
#define PICUGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum pcg_status {
  PCG_OK = 0,
  PCG_ERR_INVALID = 1,    /* bad argument (null pointer, bad flag)                          */
  PCG_ERR_NOT_INIT = 2,   /* pcg_init() was not called or failed                            */
  PCG_ERR_NO_DEVICE = 3,  /* no HIP device / the device cannot be used                      */
  PCG_ERR_HIP = 4,        /* a HIP runtime call failed                                      */
  PCG_ERR_TOO_LARGE = 5,  /* operands exceed the engine limits; analog of the reference's   */
                          /* throw "FFT length is too small." (NTT.cpp:352-353)              */
  PCG_ERR_OOM = 6,        /* device or pinned-host allocation failed                        */
  PCG_ERR_INTERNAL = 7    /* an internal consistency check failed (e.g. carry overflow)     */
} pcg_status;

/* Engine flags (the reference picks FFT for k < 27, NTT otherwise,
 * BigFloat.cpp:386; here AUTO picks per size class). */
#define PCG_ENGINE_AUTO 0u
#define PCG_ENGINE_NTT 1u  /* two-prime NTT, always exact                            */
#define PCG_ENGINE_FFT 2u  /* FP64 complex FFT with round-off check and NTT recompute */
#define PCG_ENGINE_MASK 3u

/* Limits: na + nb <= 2^30 limbs and min(na, nb) <= 2^29 limbs. */
#define PCG_MAX_TOTAL_LIMBS (1ull << 30)
#define PCG_MAX_MIN_LIMBS (1ull << 29)

/* Initialise on HIP device `device` (tables, stream).  Replaces the table
 * pre-sizing of Pi() (PI.h:145-146) and NNT::init / fft_ensure_table
 * (NTT.cpp:169-188,297-304; FFT.cpp:31-60).  Idempotent. */
int pcg_init(int device);

/* Release device memory and the stream. */
void pcg_shutdown(void);

/* Text of the last error on the calling thread ("" if none). */
const char* pcg_last_error_string(void);

/* Make sure scratch for products up to `total_limbs` = na + nb is allocated
 * (the analog of ntt_ensure_table(k) / fft_ensure_table(k), NTT.h:4, FFT.h:4). */
int pcg_reserve(uint64_t total_limbs);

/* out[0 .. na+nb) = a[0 .. na) * b[0 .. nb), host pointers, untrimmed
 * (out may have one zero top limb; the caller trims as BigFloat.cpp:411-412).
 * Replaces the engine block of BigFloat::mul, BigFloat.cpp:377-441.
 * na == 0 or nb == 0 writes na+nb zero limbs. */
int pcg_mul(const uint32_t* a, uint64_t na, const uint32_t* b, uint64_t nb, uint32_t* out,
            uint32_t flags);

/* Same on device pointers, enqueued on the library stream; returns after
 * enqueueing unless PCG_SYNC is set in flags. */
#define PCG_SYNC 0x100u
int pcg_mul_dev(const uint32_t* d_a, uint64_t na, const uint32_t* d_b, uint64_t nb,
                uint32_t* d_out, uint32_t flags);

/* pcg_mul_dev for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  The same exact product on the NTT engine
 * (PCG_ENGINE_AUTO or PCG_ENGINE_NTT; PCG_SYNC as for pcg_mul_dev), with
 * every scratch buffer carved from one static device array of 32.1 MiB that
 * each GPU holds from library load to process end: the call makes no
 * hipMalloc, hipFree or other allocation.  Limit: na + nb <= 2^20 limbs,
 * else PCG_ERR_TOO_LARGE (nothing is launched).  All calls on a GPU share
 * its array and leave their transforms in it; calls that overlap on one GPU
 * (e.g. graph replays on different streams) race. */
#define PCG_NOALLOC_MAX_LIMBS (1ull << 20)
int pcg_mul_dev_noalloc(const uint32_t* d_a, uint64_t na, const uint32_t* d_b, uint64_t nb,
                        uint32_t* d_out, uint32_t flags);

/* One product of a batch. */
typedef struct pcg_mul_desc {
  const uint32_t* a;
  uint64_t na;
  const uint32_t* b;
  uint64_t nb;
  uint32_t* out; /* na + nb limbs */
} pcg_mul_desc;

/* A batch of independent products (host pointers).  Replaces `count`
 * calls of the BigFloat::mul engine block (BigFloat.cpp:377-441) made by one
 * level of the binary-splitting tree (PI.h:110-112): products are grouped
 * into size classes and each class runs as one batched launch sequence.
 * PCG_BATCH_PLAIN instead runs one launch sequence per product (the plain
 * method, for comparison). */
#define PCG_BATCH_PLAIN 0x200u
int pcg_mul_batch(const pcg_mul_desc* descs, uint64_t count, uint32_t flags);

/* Same with device pointers in the descriptors (descs itself is host memory). */
int pcg_mul_batch_dev(const pcg_mul_desc* descs, uint64_t count, uint32_t flags);

/* Wait for all work enqueued on the library stream. */
int pcg_sync(void);

/* Device memory helpers for callers without their own allocator. */
int pcg_dev_alloc(void** ptr, uint64_t bytes);
int pcg_dev_free(void* ptr);
/* Page-locked host memory (hipHostMalloc) for large results such as the hex
 * text of pcg_df_to_hex, which then arrives at full PCIe bandwidth. */
int pcg_host_alloc(void** ptr, uint64_t bytes);
int pcg_host_free(void* ptr);
int pcg_memcpy_h2d(void* dst, const void* src, uint64_t bytes);
int pcg_memcpy_d2h(void* dst, const void* src, uint64_t bytes);

/* FFT-engine round-off report of the last PCG_ENGINE_FFT product:
 * max distance to the nearest integer over all output coefficients and
 * whether the product was recomputed on the NTT (distance > 0.25). */
int pcg_fft_last_roundoff(double* max_dist, int* recomputed);

/* Timing of the last product run with PCG_PROFILE in flags: per-stage
 * kernel durations in milliseconds measured with HIP events on the library
 * stream.  stage_ms[0] forward(a), [1] forward(b)+pointwise+inverse,
 * [2] carry; total_ms = whole enqueued sequence. */
#define PCG_PROFILE 0x400u
int pcg_last_profile(double* stage_ms, int n_stage, double* total_ms);

/* One value of a binary-splitting subtree, written into caller memory. */
typedef struct pcg_bsr_value {
  uint32_t* limbs;  /* caller buffer of `cap` limbs                         */
  uint64_t cap;
  uint64_t L;       /* limbs written (0 = zero)                              */
  int32_t sign;     /* 1 positive or zero, 0 negative (BigFloat::sign)        */
} pcg_bsr_value;

/* P, Q, R of the Chudnovsky binary-splitting subtree [a, b) (leaves and
 * combine of PI.h:64-112, midpoints m = (a + b) / 2) computed entirely on
 * the GPU.  Valid when every value of the subtree has at most max_limbs =
 * p + 2 limbs: BigFloat then never truncates, every value is an exact
 * integer (exponent 0), and the results equal those of the BigFloat
 * recursion bit for bit.  Returns PCG_ERR_TOO_LARGE when a length bound
 * exceeds max_limbs or a buffer is too small (nothing is written then).
 * pcg_bsr_bound gives the limb bound of P, Q and R for sizing buffers (from
 * the bit lengths of the leaf terms at b, which bound every leaf of [a, b)). */
int pcg_bsr_exact(uint32_t a, uint32_t b, uint64_t max_limbs, pcg_bsr_value* P, pcg_bsr_value* Q,
                  pcg_bsr_value* R);
int pcg_bsr_bound(uint32_t a, uint32_t b, uint64_t* bound);
/* Builds the evaluation plans of subtrees of the given sizes (b - a; the
 * first two distinct ones, as many as the plan cache holds): arena offsets,
 * product and add lists, laid out on the GPU from a table of the subtree
 * sizes.  pcg_bsr_exact and pcg_df_bsr_exact build a missing plan
 * themselves; calling this first moves the plan allocation ahead of the
 * first evaluation.  The reference has no plans (Pi_BSR recurses,
 * PI.h:64-112). */
int pcg_bsr_prepare(const uint32_t* sizes, uint64_t n);
/* Selects how pcg_bsr_exact and pcg_df_bsr_exact run a subtree's level
 * schedule: enable = 1 records it once per plan as a HIP graph (captured
 * on the plan's first evaluation, again after the batched scratch it names
 * was reallocated) and replays the graph for every subtree of that plan;
 * enable = 0 (the default) launches it kernel by kernel.  The results are
 * the same.  The reference recurses instead (Pi_BSR, PI.h:93-112, one
 * std::thread per branch); no counterpart there. */
int pcg_bsr_graphs(int enable);
/* Self-check, host only (no GPU call, no pcg_init needed): lays out the plan
 * of a subtree of `leaves` leaves whose largest leaf index is bmax (the
 * plans size every value from the leaf bounds at bmax; 0 = any leaf) as the
 * GPU does (the same per-node code, run on the CPU) with the small-product
 * class table `table` (n entries, as pcg_batch_table) and compares every
 * entry with a direct node-by-node construction.  PCG_OK when they agree,
 * PCG_ERR_INTERNAL naming the first difference otherwise.  No counterpart in
 * the reference. */
int pcg_bsr_plan_check(uint32_t leaves, uint32_t bmax, const int* table, int n);

/* ---- Device-resident BigFloat values ----
 * A BigFloat (BigFloat.h:33-36) whose limbs stay in device memory, so the
 * top of the binary splitting, the Newton division and 1/sqrt(10005) of
 * Pi() (PI.h:175-209) run without moving values through host memory.
 * Value = sum limbs[i] 2^(32 (exp + i)); zero has L = 0.  Limb buffers come
 * from the library's pool: start from a zeroed struct, let the calls below
 * (re)fill it, release it with pcg_df_free before pcg_shutdown.  An output
 * may be the same struct as an input.  Each call waits for its result (the
 * new length depends on the limbs), so metadata is valid on return. */
typedef struct pcg_dfloat {
  uint32_t* limbs;  /* device memory owned by the library pool              */
  uint64_t cap;     /* limbs allocated                                        */
  uint64_t L;       /* limbs in use (0 = zero)                                */
  int64_t exp;      /* exponent in limbs (BigFloat::exp)                      */
  int32_t sign;     /* 1 positive or zero, 0 negative (BigFloat::sign)        */
} pcg_dfloat;

/* Returns x's buffer to the pool; x becomes zero. */
int pcg_df_free(pcg_dfloat* x);
/* x = (sign, exp, limbs[0..L)) from host memory (BigFloat(T, L, exp, sign)). */
int pcg_df_set(pcg_dfloat* x, const uint32_t* limbs, uint64_t L, int64_t exp, int32_t sign);
/* Copies limbs [first, first + count) of x to host memory. */
int pcg_df_read(const pcg_dfloat* x, uint64_t first, uint64_t count, uint32_t* out);
/* z = x (BigFloat copy constructor, BigFloat.cpp:24-33). */
int pcg_df_copy(pcg_dfloat* z, const pcg_dfloat* x);
/* BigFloat::ucmp (BigFloat.cpp:123-147): -1, 0, 1 for |a| <, =, > |b|. */
int pcg_df_ucmp(const pcg_dfloat* a, const pcg_dfloat* b, int* cmp);
/* z = a.add(b, p) / a.sub(b, p) (BigFloat.cpp:288-325 with uadd/usub
 * 184-287): aligned by exponent, the bottom cut to p + 2 limbs (p = 0:
 * exact), a carry limb appended, or every leading zero limb stripped. */
int pcg_df_add(pcg_dfloat* z, const pcg_dfloat* a, const pcg_dfloat* b, uint64_t p);
int pcg_df_sub(pcg_dfloat* z, const pcg_dfloat* a, const pcg_dfloat* b, uint64_t p);
/* z = a.mul(b, p) (BigFloat.cpp:326-444): operands cut to their top p + 2
 * limbs (p = 0: exact), exact product on the GPU engine chosen by `flags`
 * (PCG_ENGINE_*), one zero top limb trimmed.  PCG_KEEP_B in flags keeps b's
 * forward transform for the next product; PCG_REUSE_B says that b is the
 * operand kept that way (the same pcg_dfloat, unchanged, and the same p):
 * its transform is reused when nothing has overwritten it and the transform
 * length agrees, else recomputed.  Results are identical either way.  The
 * Newton steps of BigFloat.cpp:445-567 multiply r0 twice in a row. */
#define PCG_KEEP_B 0x1000u
#define PCG_REUSE_B 0x2000u
int pcg_df_mul(pcg_dfloat* z, const pcg_dfloat* a, const pcg_dfloat* b, uint64_t p, uint32_t flags);
/* z1 = x1.mul(s, p) and z2 = x2.mul(s, p), each exactly as pcg_df_mul, with
 * the forward transform of the shared operand computed once when both
 * products have the same transform length. Replaces the two products of
 * one node combine that share an operand (PI.h:110-112: Q1 in P0*Q1 and
 * Q0*Q1, R0 in P1*R0 and R0*R1); the reference transforms it twice
 * (BigFloat.cpp:377-441 per call). z1 and z2 must be distinct; either may be
 * one of the operands. */
int pcg_df_mul_pair(pcg_dfloat* z1, const pcg_dfloat* x1, pcg_dfloat* z2, const pcg_dfloat* x2,
                    const pcg_dfloat* s, uint64_t p, uint32_t flags);
/* z = a.mul(x) (BigFloat.cpp:156-183). */
int pcg_df_mul_u32(pcg_dfloat* z, const pcg_dfloat* a, uint32_t x);
/* to_string_hex (BigFloat.cpp:95-106): the top limb in decimal, '.', then
 * 8 hex digits per lower limb, cut to 2 + digits characters; NUL-terminated
 * in out[0..cap); *written = characters before the NUL.  PCG_ERR_INVALID
 * ("Limit Exceeded") when the value has too few limbs. */
int pcg_df_to_hex(const pcg_dfloat* x, uint64_t digits, char* out, uint64_t cap, uint64_t* written);
/* pcg_bsr_exact with the results left in device memory as values of
 * exponent 0 (no copy through the host). */
int pcg_df_bsr_exact(uint32_t a, uint32_t b, uint64_t max_limbs, pcg_dfloat* P, pcg_dfloat* Q,
                     pcg_dfloat* R);
/* Frees the pool's cached buffers (values in use are untouched). */
int pcg_df_release_pool(void);

/* Size classes of batched products (pcg_mul_batch*, pcg_bsr_exact): a
 * product of zL = na + nb limbs with 2^(b-1) < zL <= 2^b, b <= 12, runs in
 * class table[b]: 0 the schoolbook kernel (zL <= 1024), 1..3 the grouped
 * schoolbook (16 threads x 4 columns, zL <= 64; 64 x 4, zL <= 256; 64 x 8,
 * zL <= 512), 4 one Karatsuba level over the schoolbook (zL <= 1024),
 * 10..12 the single-workgroup NTT of 2^k points, 110..112 the
 * batched pass kernels at 2^(k-100) points.  Entries 13..20 (when n > 13)
 * are the pass geometry of batched products at 2^b points: 0 the
 * 4096-point passes, 1 the large-tile passes (2^14-point tiles; at 2^13
 * points the one-pass 2^14-point product).  The crossovers are the
 * autotuner's (timed on the device
 * at the first batched call, or by pcg_autotune, which returns the table);
 * pcg_batch_table reads it (set = 0) or installs one (set = 1, every entry
 * valid for its bucket).  Products are exact in every class and geometry;
 * the table only moves work between kernels.  The reference picks its
 * engine by size alone (BigFloat.cpp:386). */
int pcg_autotune(int* table, int n);
int pcg_batch_table(int* table, int n, int set);

/* Elapsed time of the work enqueued on the library stream between the two
 * calls, measured with HIP events recorded on that stream. */
int pcg_timer_start(void);
int pcg_timer_stop(double* ms);

/* Static facts for reports: transform length (log2) chosen for na + nb
 * limbs by the NTT engine, and the number of passes. */
int pcg_ntt_plan(uint64_t total_limbs, int* log2_points, int* passes);

#ifdef __cplusplus
}
#endif
#endif
