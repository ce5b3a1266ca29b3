// Chudnovsky binary splitting on the GPU BigFloat (see Pi.h).
#include "Pi.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/picugpu.h"
#include "BigFloat.h"
#include "DeviceFloat.h"

namespace piculator {

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct PQR {
  BigFloat P, Q, R;
};

// Leaf terms of PI.h:67-88:
//   P = -(13591409 + 545140134 b)(2b-1)(6b-5)(6b-1) for odd b (sign (-1)^b)
//   Q = 10939058860032000 b^3 = b^3 * 26726400 * 409297880
//   R = (2b-1)(6b-5)(6b-1)
void leaf(uint32_t b, PQR& out) {
  BigFloat P = BigFloat(b).mul(545140134u);
  P = P.add(BigFloat(13591409u));
  P = P.mul(2 * b - 1);
  P = P.mul(6 * b - 5);
  P = P.mul(6 * b - 1);
  if (b % 2 == 1) P.negate();
  BigFloat Q = BigFloat(b).mul(b).mul(b).mul(26726400u).mul(409297880u);
  BigFloat R = BigFloat(2 * b - 1).mul(6 * b - 5).mul(6 * b - 1);
  out.P = std::move(P);
  out.Q = std::move(Q);
  out.R = std::move(R);
}

// PI.h:64-112, depth first; each mul is its own GPU call.
void bsr_plain(PQR& out, uint32_t a, uint32_t b, size_t p) {
  if (b - a == 1) {
    leaf(b, out);
    return;
  }
  const uint32_t m = (a + b) / 2;
  PQR L, R;
  bsr_plain(L, a, m, p);
  bsr_plain(R, m, b, p);
  out.P = L.P.mul(R.Q, p).add(R.P.mul(L.R, p), p);
  out.Q = L.Q.mul(R.Q, p);
  out.R = L.R.mul(R.R, p);
}

// ---------------------------------------------------------------- batched
struct Node {
  uint32_t a, b;
  int32_t left, right;  // child node indices, -1 for a leaf
  int32_t height;       // 1 for leaves
};

// Explicit tree of [a, b) with the reference's midpoints (PI.h:90).
int32_t build(std::vector<Node>& nodes, uint32_t a, uint32_t b) {
  const int32_t id = (int32_t)nodes.size();
  nodes.push_back(Node{a, b, -1, -1, 1});
  if (b - a > 1) {
    const uint32_t m = (a + b) / 2;
    const int32_t l = build(nodes, a, m);
    const int32_t r = build(nodes, m, b);
    nodes[id].left = l;
    nodes[id].right = r;
    nodes[id].height = std::max(nodes[l].height, nodes[r].height) + 1;
  }
  return id;
}

// Runs body(i) for i in [0, n) on the host threads (independent items).
template <class F>
void parallel_for(size_t n, F body) {
  unsigned T = std::thread::hardware_concurrency();
  if (T > 16) T = 16;
  if (T <= 1 || n < 256) {
    for (size_t i = 0; i < n; ++i) body(i);
    return;
  }
  std::vector<std::thread> th;
  const size_t per = (n + T - 1) / T;
  for (unsigned t = 0; t < T; ++t) {
    const size_t lo = t * per, hi = std::min(n, lo + per);
    if (lo >= hi) break;
    th.emplace_back([lo, hi, &body]() {
      for (size_t i = lo; i < hi; ++i) body(i);
    });
  }
  for (std::thread& x : th) x.join();
}

// Evaluates the subtree [a, b) bottom-up on host BigFloats, batching every
// height's products (MulBatch); the fallback for the small subtrees whose
// values exceed the precision window at very low digit counts.
void bsr_subtree_batched(PQR& out, uint32_t a, uint32_t b, size_t p, bool verbose) {
  std::vector<Node> nodes;
  nodes.reserve(2 * (size_t)(b - a));
  const int32_t root = build(nodes, a, b);
  std::vector<PQR> val(nodes.size());
  const int32_t H = nodes[root].height;
  std::vector<std::vector<int32_t>> by_h(H + 1);
  for (int32_t i = 0; i < (int32_t)nodes.size(); ++i) by_h[nodes[i].height].push_back(i);
  double tl = now_s();
  {
    const std::vector<int32_t>& leaves = by_h[1];
    parallel_for(leaves.size(), [&](size_t k) { leaf(nodes[leaves[k]].b, val[leaves[k]]); });
  }
  if (verbose) std::fprintf(stderr, "  leaves: %zu, %.3f s\n", by_h[1].size(), now_s() - tl);
  std::vector<BigFloat> x, y;
  for (int32_t h = 2; h <= H; ++h) {
    const std::vector<int32_t>& ids = by_h[h];
    x.resize(ids.size());
    y.resize(ids.size());
    const double ta = now_s();
    MulBatch mb;
    for (size_t k = 0; k < ids.size(); ++k) {
      const Node& n = nodes[ids[k]];
      PQR& L = val[n.left];
      PQR& R = val[n.right];
      PQR& O = val[ids[k]];
      mb.add(L.P, R.Q, p, &x[k]);
      mb.add(R.P, L.R, p, &y[k]);
      mb.add(L.Q, R.Q, p, &O.Q);
      mb.add(L.R, R.R, p, &O.R);
    }
    const double t0 = now_s();
    mb.run(false);
    const double t1 = now_s();
    parallel_for(ids.size(), [&](size_t k) {
      const Node& n = nodes[ids[k]];
      val[ids[k]].P = x[k].add(y[k], p);
      val[n.left] = PQR();
      val[n.right] = PQR();
    });
    if (verbose)
      std::fprintf(stderr, "  height %d: %zu nodes, queue %.3f s, products %.3f s, adds %.3f s\n", h,
                   ids.size(), t0 - ta, t1 - t0, now_s() - t1);
  }
  out = std::move(val[root]);
}

// Largest subtree handed to the GPU-resident evaluation at once.  Each
// level of such a subtree runs as one batched launch set, where the combines
// above it run node by node, so larger is faster while the plans fit: at
// 10^9 digits this gives 32 subtrees of 2.2 M leaves (4.1 GB of arena per
// plan, two plans).
constexpr uint32_t kDeviceLeaves = 1u << 22;

// Values of a node kept in device memory.
struct DPQR {
  DeviceFloat P, Q, R;
};

// Subtree [a, b) evaluated on the GPU, results left in device memory; false
// when a value of it could exceed p + 2 limbs.
bool bsr_device_resident(DPQR& out, uint32_t a, uint32_t b, size_t p) {
  uint64_t bound = 0;
  if (pcg_bsr_bound(a, b, &bound) != PCG_OK || bound > p + YCL_BIGFLOAT_EXTRA_PRECISION) return false;
  const uint64_t mx = p + YCL_BIGFLOAT_EXTRA_PRECISION;
  int rc = pcg_df_bsr_exact(a, b, mx, &out.P.raw(), &out.Q.raw(), &out.R.raw());
  if (rc == PCG_ERR_NOT_INIT && pcg_init(0) == PCG_OK)
    rc = pcg_df_bsr_exact(a, b, mx, &out.P.raw(), &out.Q.raw(), &out.R.raw());
  if (rc == PCG_ERR_TOO_LARGE) return false;
  if (rc != PCG_OK) throw pcg_last_error_string();
  return true;
}

// Distinct sizes of the subtrees bsr_resident hands to the GPU (the same
// walk and the same bound test as bsr_device_resident).
void resident_sizes(uint32_t a, uint32_t b, size_t p, std::vector<uint32_t>& sizes) {
  if (b - a <= kDeviceLeaves) {
    uint64_t bound = 0;
    if (pcg_bsr_bound(a, b, &bound) == PCG_OK && bound <= p + YCL_BIGFLOAT_EXTRA_PRECISION) {
      if (std::find(sizes.begin(), sizes.end(), b - a) == sizes.end()) sizes.push_back(b - a);
      return;
    }
  }
  if (b - a <= 64) return;
  const uint32_t m = (a + b) / 2;
  resident_sizes(a, m, p, sizes);
  resident_sizes(m, b, p, sizes);
}

// The series with every value resident in device memory: GPU subtrees
// below, the node combine of PI.h:95-110 on DeviceFloat above (the same
// truncation as BigFloat, so the same values).
void bsr_resident(DPQR& out, uint32_t a, uint32_t b, size_t p, bool verbose) {
  const double t0 = verbose ? now_s() : 0.0;
  if (b - a <= kDeviceLeaves && bsr_device_resident(out, a, b, p)) {
    if (verbose)
      std::fprintf(stderr, "  device [%u, %u): %.3f s, |P| %zu |Q| %zu |R| %zu limbs\n", a, b,
                   now_s() - t0, out.P.get_precision(), out.Q.get_precision(),
                   out.R.get_precision());
    return;
  }
  if (b - a <= 64) {
    PQR h;
    bsr_subtree_batched(h, a, b, p, false);
    out.P = DeviceFloat(h.P);
    out.Q = DeviceFloat(h.Q);
    out.R = DeviceFloat(h.R);
    return;
  }
  const uint32_t m = (a + b) / 2;
  DPQR L, R;
  bsr_resident(L, a, m, p, verbose);
  bsr_resident(R, m, b, p, verbose);
  const double t1 = verbose ? now_s() : 0.0;
  // P0 Q1 and Q0 Q1 share the transform of Q1, P1 R0 and R0 R1 that of R0
  auto xq = R.Q.mul_pair(L.P, L.Q, p);  // x = P0 Q1, Q = Q0 Q1
  auto yr = L.R.mul_pair(R.P, R.R, p);  // y = P1 R0, R = R0 R1
  out.Q = std::move(xq.second);
  out.R = std::move(yr.second);
  const double t2 = verbose ? now_s() : 0.0;
  out.P = xq.first.add(yr.first, p);
  if (verbose)
    std::fprintf(stderr, "  combine [%u, %u): products %.3f s, add %.3f s, |Q| %zu limbs\n", a, b,
                 t2 - t1, now_s() - t2, out.Q.get_precision());
}

}  // namespace

void pi_sizes(size_t decimal_digits, size_t* hex_digits, size_t* p, size_t* terms) {
  const size_t hex = (size_t)(decimal_digits * std::log(10) / std::log(16));
  const size_t pp = (hex + 9) / 9;
  const size_t p_dec = (decimal_digits + 9) / 9;
  const size_t t = (size_t)(p_dec * 0.6346230241342037371474889163921741077188431452678) + 1;
  if (hex_digits) *hex_digits = hex;
  if (p) *p = pp;
  if (terms) *terms = t;
}

size_t pi_hex_into(size_t decimal_digits, PiMode mode, char* out, size_t cap, PiTimes* times,
                   bool verbose) {
  size_t hex, p, terms;
  pi_sizes(decimal_digits, &hex, &p, &terms);
  if ((uint32_t)terms != terms) throw "Limit Exceeded";
  const uint64_t m0 = big_mul_count();
  const double t0 = now_s();
  PiTimes pt;  // phase durations
  size_t n = 0;
  if (mode == PiMode::Plain) {
    PQR s;
    bsr_plain(s, 0, (uint32_t)terms, p);
    BigFloat P = s.Q.mul(13591409u).add(s.P, p);
    BigFloat Q = s.Q.mul(4270934400u);
    const double t1 = now_s();
    pt.series = t1 - t0;
    P = Q.div(P, p);
    const double t2 = now_s();
    pt.division = t2 - t1;
    Q = invsqrt(10005, p);
    const double t3 = now_s();
    pt.invsqrt = t3 - t2;
    P = P.mul(Q, p);
    const double t4 = now_s();
    pt.final_mul = t4 - t3;
    const std::string str = P.to_string_hex(hex);
    if (str.size() + 1 > cap) throw "output capacity too small";
    std::memcpy(out, str.c_str(), str.size() + 1);
    n = str.size();
    pt.conversion = now_s() - t4;
  } else {
    // The subtree plans (laid out on the GPU), then 1/sqrt(10005): it is
    // independent of the series, so computing it first changes no value.
    std::vector<uint32_t> sizes;
    resident_sizes(0, (uint32_t)terms, p, sizes);
    if (!sizes.empty() && pcg_bsr_prepare(sizes.data(), sizes.size()) != PCG_OK)
      throw pcg_last_error_string();
    const double ti = now_s();
    const DeviceFloat S = invsqrt_device(10005, p);
    const double tb = now_s();
    pt.invsqrt = tb - ti;
    DPQR s;
    bsr_resident(s, 0, (uint32_t)terms, p, verbose);
    DeviceFloat P = s.Q.mul(13591409u).add(s.P, p);
    const DeviceFloat Q = s.Q.mul(4270934400u);
    s = DPQR();
    const double t1 = now_s();
    pt.series = t1 - tb;
    P = Q.div(P, p);
    const double t2 = now_s();
    pt.division = t2 - t1;
    P = P.mul(S, p);
    const double t4 = now_s();
    pt.final_mul = t4 - t2;
    n = P.to_hex_into(hex, out, cap);
    pt.conversion = now_s() - t4;
  }
  if (times) {
    *times = pt;
    times->total = now_s() - t0;
    times->big_muls = big_mul_count() - m0;
  }
  return n;
}

std::string pi_hex(size_t decimal_digits, PiMode mode, PiTimes* times, bool verbose) {
  size_t hex, p, terms;
  pi_sizes(decimal_digits, &hex, &p, &terms);
  std::string s(hex + 32, '\0');
  s.resize(pi_hex_into(decimal_digits, mode, &s[0], s.size(), times, verbose));
  return s;
}

void Pi(size_t decimal_digits, size_t threads) {
  size_t hex, p, terms;
  pi_sizes(decimal_digits, &hex, &p, &terms);
  std::printf("Constant:\tPi\nAlgorithm:\tChudnovsky (1988)\n\n");
  std::printf("Decimal Digits:\t%zu\nHex Digits:\t%zu\n\n", decimal_digits, hex);
  std::printf("Memory Mode:\tGPU (MI355X HBM)\n# Threads:\t%zu (host; products run on the GPU)\n\n",
              threads);
  std::printf("Summing Series...  %zu\n", terms);
  PiTimes t;
  const std::string str = pi_hex(decimal_digits, PiMode::Batched, &t);
  std::printf("series %.3f s\nDivision... %.3f s\nInvSqrt... %.3f s\nFinal Multiply... %.3f s\n"
              "Base Conversion... %.3f s\nTotal Time:\t%.3f s\n",
              t.series, t.division, t.invsqrt, t.final_mul, t.conversion, t.total);
  FILE* f = std::fopen("pi.txt", "wb");
  if (!f) throw "Cannot Create File";
  std::fwrite(str.data(), 1, str.size(), f);
  std::fclose(f);
}

}  // namespace piculator
