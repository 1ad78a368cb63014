/*
 * qdb_gen.h - seeded synthetic inputs for the aggregate and join workloads.
 *
 * Counter-based: value (seed, stream, row) depends on nothing else, so a host
 * (C) and a device (HIP) build produce identical bits for any row range, and
 * a GPU can generate 100M rows in HBM while a CPU checker regenerates any
 * slice of them. Only integer arithmetic and IEEE add/multiply in a fixed
 * order (no fused multiply-add) are used.
 *
 * Workloads (sizes are the caller's):
 *   lineitem  flag i8 U{0,1,2}; status i8 U{0,1}; shipdate i32 U[0,2525];
 *             quantity DECIMAL(15,2) = 100*U{1..50}; price DECIMAL(15,2)
 *             U[90000, 10500000]; discount DECIMAL(15,2) U{0..10};
 *             tax DECIMAL(15,2) U{0..8}.
 *   join      build key i64 = a seeded permutation of 1..B, segment i32 U{0..4};
 *             probe fk i64 U{1..B}, value DECIMAL(15,2) U[90000, 10500000].
 *   groups    key i64 = a bijective mix of a Zipf(s) rank over D values,
 *             vi i64 U[-1e6, 1e6], vf f64 ~ N(0, 1e3) (Irwin-Hall, 12 terms).
 */
#ifndef QDB_GEN_H
#define QDB_GEN_H

#include <stdint.h>

#if defined(__HIPCC__)
#define QDB_GEN_FN static inline __host__ __device__
#else
#define QDB_GEN_FN static inline
#endif

enum {
    QDB_STREAM_FLAG = 1, QDB_STREAM_STATUS, QDB_STREAM_SHIPDATE, QDB_STREAM_QUANTITY,
    QDB_STREAM_PRICE, QDB_STREAM_DISCOUNT, QDB_STREAM_TAX,
    QDB_STREAM_BUILD_PERM, QDB_STREAM_SEGMENT, QDB_STREAM_FK, QDB_STREAM_VALUE,
    QDB_STREAM_ZIPF, QDB_STREAM_VI, QDB_STREAM_VF, QDB_STREAM_KEYMIX
};

/* splitmix64 finalizer: a bijection on 64-bit words. */
QDB_GEN_FN uint64_t qdb_mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

QDB_GEN_FN uint64_t qdb_rand(uint64_t seed, uint64_t stream, uint64_t row)
{
    return qdb_mix64(qdb_mix64(seed + stream * 0x9E3779B97F4A7C15ull) ^ (row * 0xD1B54A32D192ED03ull + stream));
}

/* High 64 bits of the 128-bit product a * b, from 32-bit partial products
 * (the same operations on the host and the device). */
QDB_GEN_FN uint64_t qdb_mulhi64(uint64_t a, uint64_t b)
{
    const uint64_t a_lo = a & 0xffffffffu, a_hi = a >> 32;
    const uint64_t b_lo = b & 0xffffffffu, b_hi = b >> 32;
    const uint64_t p0 = a_lo * b_lo, p1 = a_lo * b_hi, p2 = a_hi * b_lo, p3 = a_hi * b_hi;
    const uint64_t mid = (p0 >> 32) + (p1 & 0xffffffffu) + (p2 & 0xffffffffu);
    return p3 + (p1 >> 32) + (p2 >> 32) + (mid >> 32);
}

/* Uniform integer in [lo, hi]: the high word of rand * span (multiply-shift
 * range reduction, no division). A 64-bit remainder by a constant is avoided
 * on purpose: hipcc (ROCm 7.2) has lowered `x % 50` inside the lineitem
 * generator kernel for gfx950 to code that drops bits of the quotient. */
QDB_GEN_FN int64_t qdb_uniform(uint64_t seed, uint64_t stream, uint64_t row, int64_t lo, int64_t hi)
{
    uint64_t span = (uint64_t)(hi - lo) + 1u;
    return lo + (int64_t)qdb_mulhi64(qdb_rand(seed, stream, row), span);
}

/* Uniform double in [0, 1) on the 2^-53 grid. */
QDB_GEN_FN double qdb_unit(uint64_t r)
{
    return (double)(r >> 11) * (1.0 / 9007199254740992.0);
}

/* Bijection of [0, n): a 4-round Feistel network over the smallest even
 * bit-width domain >= n, with cycle walking. */
QDB_GEN_FN uint64_t qdb_permute(uint64_t seed, uint64_t x, uint64_t n)
{
    unsigned bits = 2;
    while ((1ull << bits) < n) bits += 2;
    const unsigned half = bits / 2;
    const uint64_t mask = (1ull << half) - 1u;
    do {
        uint64_t l = x >> half, r = x & mask;
        for (unsigned round = 0; round < 4; ++round) {
            uint64_t f = qdb_mix64(seed * 0x9E3779B97F4A7C15ull + r * 0xBF58476D1CE4E5B9ull + round) & mask;
            uint64_t t = l ^ f;
            l = r;
            r = t;
        }
        x = (l << half) | r;
    } while (x >= n);
    return x;
}

QDB_GEN_FN void qdb_gen_lineitem_row(uint64_t seed, int64_t row, int8_t* flag, int8_t* status,
    int32_t* shipdate, int64_t* quantity, int64_t* price, int64_t* discount, int64_t* tax)
{
    uint64_t r = (uint64_t)row;
    *flag = (int8_t)qdb_uniform(seed, QDB_STREAM_FLAG, r, 0, 2);
    *status = (int8_t)qdb_uniform(seed, QDB_STREAM_STATUS, r, 0, 1);
    *shipdate = (int32_t)qdb_uniform(seed, QDB_STREAM_SHIPDATE, r, 0, 2525);
    *quantity = 100 * qdb_uniform(seed, QDB_STREAM_QUANTITY, r, 1, 50);
    *price = qdb_uniform(seed, QDB_STREAM_PRICE, r, 90000, 10500000);
    *discount = qdb_uniform(seed, QDB_STREAM_DISCOUNT, r, 0, 10);
    *tax = qdb_uniform(seed, QDB_STREAM_TAX, r, 0, 8);
}

QDB_GEN_FN void qdb_gen_join_build_row(uint64_t seed, int64_t row, int64_t build_rows, int64_t* key, int32_t* segment)
{
    uint64_t r = (uint64_t)row;
    *key = 1 + (int64_t)qdb_permute(qdb_mix64(seed + QDB_STREAM_BUILD_PERM), r, (uint64_t)build_rows);
    *segment = (int32_t)qdb_uniform(seed, QDB_STREAM_SEGMENT, r, 0, 4);
}

QDB_GEN_FN void qdb_gen_join_probe_row(uint64_t seed, int64_t row, int64_t build_rows, int64_t* fk, int64_t* value)
{
    uint64_t r = (uint64_t)row;
    *fk = qdb_uniform(seed, QDB_STREAM_FK, r, 1, build_rows);
    *value = qdb_uniform(seed, QDB_STREAM_VALUE, r, 90000, 10500000);
}

/* First index k of cdf[0, n) with cdf[k] > u (cdf ascending, cdf[n-1] = 1). */
QDB_GEN_FN int64_t qdb_cdf_search(const double* cdf, int64_t n, double u)
{
    int64_t lo = 0, hi = n - 1;
    while (lo < hi) {
        int64_t mid = lo + ((hi - lo) >> 1);
        if (cdf[mid] > u) hi = mid; else lo = mid + 1;
    }
    return lo;
}

QDB_GEN_FN double qdb_irwin_hall_normal(uint64_t seed, uint64_t row)
{
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
    double s = 0.0;
    for (uint64_t j = 0; j < 12; ++j) {
        s = s + qdb_unit(qdb_rand(seed, QDB_STREAM_VF, row * 12u + j));
    }
    return (s - 6.0) * 1000.0;
}

QDB_GEN_FN void qdb_gen_groups_row(uint64_t seed, int64_t row, const double* cdf, int64_t distinct,
    int64_t* key, int64_t* vi, double* vf)
{
    uint64_t r = (uint64_t)row;
    double u = qdb_unit(qdb_rand(seed, QDB_STREAM_ZIPF, r));
    int64_t rank = qdb_cdf_search(cdf, distinct, u);
    *key = (int64_t)qdb_mix64((uint64_t)rank ^ qdb_mix64(seed + QDB_STREAM_KEYMIX));
    *vi = qdb_uniform(seed, QDB_STREAM_VI, r, -1000000, 1000000);
    *vf = qdb_irwin_hall_normal(seed, r);
}

/* Zipf(s = 1) cumulative distribution over ranks 1..n, summed in rank order
 * (host only; the device receives this array). */
static inline void qdb_gen_zipf1_cdf(double* cdf, int64_t n)
{
    double h = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        h += 1.0 / (double)(i + 1);
        cdf[i] = h;
    }
    for (int64_t i = 0; i < n; ++i) {
        cdf[i] = cdf[i] / h;
    }
    cdf[n - 1] = 1.0;
}

#endif /* QDB_GEN_H */
