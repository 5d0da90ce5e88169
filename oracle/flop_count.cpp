// Operation counter over the oracle's TriDist restatement (oracle_tridist.inc
// compiled with REAL = a counting number type). Test infrastructure only:
// it fixes the static per-pair FP32 operation count used in throughput
// figures; it is never part of the product path.
//
// Counting rule: add, sub, mul, div, sqrt and every comparison = 1
// (no fused operations: the reference arithmetic is compiled without FMA
// contraction). Copies and conversions of single values are free, except
// the explicit float-to-double conversion in front of the face test's
// normal-length comparison, which stands for that comparison.
#include <cmath>
#include <cstddef>
#include <cstdint>

namespace {

uint64_t g_ops = 0;

struct CF {
    double v;
    CF() : v(0) {}
    CF(double x) : v(x) {}
    explicit operator double() const {
        g_ops++;
        return v;
    }
};

inline CF operator+(CF a, CF b) { g_ops++; return CF(a.v + b.v); }
inline CF operator-(CF a, CF b) { g_ops++; return CF(a.v - b.v); }
inline CF operator*(CF a, CF b) { g_ops++; return CF(a.v * b.v); }
inline CF operator/(CF a, CF b) { g_ops++; return CF(a.v / b.v); }
inline CF operator-(CF a) { return CF(-a.v); }
inline CF &operator+=(CF &a, CF b) { g_ops++; a.v += b.v; return a; }
inline bool operator<(CF a, CF b) { g_ops++; return a.v < b.v; }
inline bool operator<=(CF a, CF b) { g_ops++; return a.v <= b.v; }
inline bool operator>(CF a, CF b) { g_ops++; return a.v > b.v; }
inline bool operator>=(CF a, CF b) { g_ops++; return a.v >= b.v; }
inline bool operator==(CF a, CF b) { g_ops++; return a.v == b.v; }
inline bool operator!=(CF a, CF b) { g_ops++; return a.v != b.v; }
inline CF cf_sqrt(CF a) { g_ops++; return CF(std::sqrt(a.v)); }

}  // namespace

#define CAT2_(a, b) a##b
#define CAT_(a, b) CAT2_(a, b)
#define REAL CF
#define SFX _c
#define SQRT cf_sqrt
#include "oracle_tridist.inc"

// operations executed by one TriDist(S, T) evaluation (9 doubles each);
// *d receives the distance, *early whether an edge-edge case returned it
// (no face test ran)
extern "C" uint64_t flop_tridist(const double *S9, const double *T9, double *d, int *early) {
    CF S[9], T[9];
    for (int i = 0; i < 9; i++) {
        S[i] = CF(S9[i]);
        T[i] = CF(T9[i]);
    }
    g_ops = 0;
    const CF r = tridist_c(S, T);
    const uint64_t n = g_ops;
    *d = r.v;
    int disjoint = 0, closest = 0;
    (void)tri_seg_c(S, T, &disjoint, &closest);
    *early = closest;
    return n;
}
