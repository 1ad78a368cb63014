/*
 * oracle.c - CPU restatement of the reference's hash-aggregate / hash-join
 * path. TEST INFRASTRUCTURE ONLY: the parity tests, the smoke check and the
 * CPU-baseline leg of bench.py load it as the checker. The product (the HIP
 * library under qumirdb_amd/) never links, loads or calls it.
 *
 * What each part restates (reference file:line):
 *   expressions   filter GenFilterKernelAst (qdb/kernel/gen.cpp:911-1091),
 *                 project GenProjectKernelAst (gen.cpp:1373-1613), decimal
 *                 product/sum typing (qdb/kernel/annotate_type.cpp:746-760,
 *                 912-960), qdb_decimal_mul (qdb/modules/qumirdb.oz:1313),
 *                 SQL three-valued logic (qdb_is_true).
 *   key hash      HashKeyValue (gen.cpp:233-379): integer mix, struct combine,
 *                 canonical float bits (gen.cpp:192-231).
 *   swiss table   rh_lookup_or_empty_physical (qdb/kernel/aggregation/
 *                 robin_hood_dual_key.oz:14-52), 8-slot SWAR groups with
 *                 triangular steps (swiss_group.oz:1-67), grow before 7/8
 *                 load (robin_hood_dual_key.oz:165-167), dense ids in
 *                 insertion order (aht_insert_dual :170-224).
 *   reducers      GenReducerFunDecls (gen.cpp:3301-3592): COUNT +1, integer
 *                 SUM i64 wrap, decimal SUM as i128 (BinInt, compiler.cpp:
 *                 2665-2683), MIN/MAX with is_new seeding, f64 SUM IEEE add
 *                 in arrival order, nullable args with a valid counter.
 *   result types  AggregateResultType (qdb/plan/ops/aggregate.cpp:21-61).
 *   finalize      agg_finish_rowset (gen.cpp:2836-3132): dense order.
 *   decimal->f64  qdb_decimal_to_f64 (qdb/modules/qumirdb.oz:983-1046) and
 *                 qdb_decimal64_to_f64 (qumirdb.oz:1224-1230).
 *   row join      rt_init / rt_mix / rt_salt / rt_seek (qdb/kernel/row_table.oz:
 *                 140-295), jr_build_rows (qdb/kernel/join/join_rows.oz:70-113),
 *                 jr_probe_candidates (join_rows.oz:146-197), append
 *                 (qdb/kernel/join_gen.cpp:2063-2159), probe pass 1
 *                 (join_gen.cpp:2161-2270), pair encoding jr_push_pair
 *                 (join_rows.oz:117-125).
 *
 * Besides the arrival-order f64 SUM of the reference, two more summations
 * are offered as checkers: Kahan-Babuska-Neumaier compensated summation, and
 * the exact sum rounded once to nearest-even (a fixed-point superaccumulator
 * over the whole double range).
 */
#include "../include/qdb_gpu.h"

#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef __int128 i128;
typedef unsigned __int128 u128;

/* ------------------------------------------------------------------------ */
/* Column access                                                             */
/* ------------------------------------------------------------------------ */

static int bit_at(const uint8_t* bitmap, int64_t index, int64_t offset)
{
    int64_t b = index + offset; /* bitoff, qdb/modules/qumirdb.oz:109-116 */
    return (bitmap[b >> 3] >> (b & 7)) & 1;
}

static int col_valid(const TColumn* c, int64_t row)
{
    return c->Mask == NULL || bit_at(c->Mask, row, c->MaskBitOffset);
}

/* Reads an integer-like column value widened to i128 (sign-extended unless
 * the kind is unsigned). */
static i128 col_int(const TColumn* c, const QdbValueType* t, int64_t row)
{
    const char* d = c->Data;
    if (t->Kind == QDB_KIND_BOOL) {
        return bit_at((const uint8_t*)d, row, c->DataBitOffset);
    }
    int uns = t->Kind == QDB_KIND_UNSIGNED;
    switch (t->Width) {
    case 1: return uns ? (i128)((const uint8_t*)d)[row] : (i128)((const int8_t*)d)[row];
    case 2: return uns ? (i128)((const uint16_t*)d)[row] : (i128)((const int16_t*)d)[row];
    case 4: return uns ? (i128)((const uint32_t*)d)[row] : (i128)((const int32_t*)d)[row];
    case 8: return uns ? (i128)((const uint64_t*)d)[row] : (i128)((const int64_t*)d)[row];
    case 16: {
        i128 v;
        memcpy(&v, d + 16 * row, 16);
        return v;
    }
    default: return 0;
    }
}

static double col_f64(const TColumn* c, const QdbValueType* t, int64_t row)
{
    if (t->Width == 4) return (double)((const float*)c->Data)[row];
    return ((const double*)c->Data)[row];
}

/* ------------------------------------------------------------------------ */
/* Decimal -> double, qdb/modules/qumirdb.oz:983-1046 (nearest, ties even). */
/* ------------------------------------------------------------------------ */

static i128 pow10_i128(int64_t scale)
{
    i128 p = 1;
    for (int64_t i = 0; i < scale; ++i) p *= 10;
    return p;
}

double oracle_decimal_to_f64(int64_t lo, int64_t hi, int64_t scale)
{
    i128 v = (i128)(((u128)(uint64_t)hi << 64) | (u128)(uint64_t)lo);
    i128 p = pow10_i128(scale);
    int negative = v < 0;
    i128 mag = negative ? -v : v;
    double result = 0.0;
    if (mag < (i128)9007199254740992LL && scale <= 22) {
        result = (double)mag / (double)p;
    } else if (mag != 0) {
        i128 q = mag / p;
        i128 r = mag - q * p;
        int64_t e = 0;
        while (q < (i128)18014398509481984LL) {
            q = q * 2;
            if (r >= p - r) {
                q = q + 1;
                r = r - (p - r);
            } else {
                r = r + r;
            }
            e = e - 1;
        }
        int guard = 0, low = r != 0;
        while (q >= (i128)9007199254740992LL) {
            if (guard) low = 1;
            guard = (int)(q & 1);
            q = q / 2;
            e = e + 1;
        }
        if (guard && (low || (q & 1))) q = q + 1;
        result = (double)q;
        while (e > 0) { result *= 2.0; --e; }
        while (e < 0) { result *= 0.5; ++e; }
    }
    return negative ? 0.0 - result : result;
}

/* qdb_decimal64_to_f64, qdb/modules/qumirdb.oz:1224-1230. */
double oracle_decimal64_to_f64(int64_t value, int64_t scale)
{
    if (value < 9007199254740992LL && value > -9007199254740992LL) {
        return (double)value / (double)(int64_t)pow10_i128(scale);
    }
    return oracle_decimal_to_f64(value, value < 0 ? -1 : 0, scale);
}

/* ------------------------------------------------------------------------ */
/* Expression programs (the spec format of include/qdb_gpu.h)                */
/* ------------------------------------------------------------------------ */

typedef struct {
    i128 i;
    double f;
    uint8_t cls;
    uint8_t valid;
} Reg;

static i128 imm128(const QdbExprInsn* in)
{
    return (i128)(((u128)(uint64_t)in->ImmHi << 64) | (u128)(uint64_t)in->ImmLo);
}

static void eval_program(const QdbExprProgram* p, const QdbColumnSchema* schema,
    const TColumn* cols, int64_t ncols, int64_t row, Reg* r)
{
    for (int k = 0; k < p->NumInsns; ++k) {
        const QdbExprInsn* in = &p->Insns[k];
        Reg* d = &r[in->Dst];
        const Reg a = r[in->A];
        const Reg b = r[in->B];
        Reg out;
        memset(&out, 0, sizeof(out));
        out.cls = in->Class;
        switch (in->Op) {
        case QDB_EXPR_COLUMN: {
            const TColumn* c = &cols[in->Column];
            const QdbValueType* t = &schema->Types[in->Column];
            out.valid = (uint8_t)col_valid(c, row);
            if (t->Kind == QDB_KIND_FLOAT) {
                out.cls = QDB_REG_F64;
                out.f = out.valid ? col_f64(c, t, row) : 0.0;
            } else {
                out.cls = t->Kind == QDB_KIND_BOOL ? QDB_REG_BOOL : QDB_REG_INT;
                out.i = out.valid ? col_int(c, t, row) : 0;
            }
            (void)ncols;
            break;
        }
        case QDB_EXPR_CONST_INT:
            out.valid = 1;
            out.i = imm128(in);
            break;
        case QDB_EXPR_CONST_F64:
            out.valid = 1;
            memcpy(&out.f, &in->ImmLo, 8);
            break;
        case QDB_EXPR_ADD:
        case QDB_EXPR_SUB:
        case QDB_EXPR_MUL:
            out.valid = a.valid && b.valid;
            if (in->Class == QDB_REG_F64) {
                double x = a.cls == QDB_REG_F64 ? a.f : (double)a.i;
                double y = b.cls == QDB_REG_F64 ? b.f : (double)b.i;
                out.f = in->Op == QDB_EXPR_ADD ? x + y : in->Op == QDB_EXPR_SUB ? x - y : x * y;
                if (!out.valid) out.f = 0.0;
            } else if (out.valid) {
                /* i128 arithmetic wraps, like the reference's BinInt ops. */
                u128 x = (u128)a.i, y = (u128)b.i;
                out.i = (i128)(in->Op == QDB_EXPR_ADD ? x + y : in->Op == QDB_EXPR_SUB ? x - y : x * y);
            }
            break;
        case QDB_EXPR_NEG:
            out.valid = a.valid;
            if (in->Class == QDB_REG_F64) out.f = a.valid ? -a.f : 0.0;
            else if (a.valid) out.i = (i128)(0 - (u128)a.i);
            break;
        case QDB_EXPR_SCALE_UP:
            out.valid = a.valid;
            if (a.valid) out.i = (i128)((u128)a.i * (u128)pow10_i128(in->ImmLo));
            break;
        case QDB_EXPR_CMP: {
            out.cls = QDB_REG_BOOL;
            out.valid = a.valid && b.valid;
            int c = 0;
            if (out.valid) {
                if (a.cls == QDB_REG_F64 || b.cls == QDB_REG_F64) {
                    double x = a.cls == QDB_REG_F64 ? a.f : (double)a.i;
                    double y = b.cls == QDB_REG_F64 ? b.f : (double)b.i;
                    switch (in->ImmLo) {
                    case QDB_CMP_EQ: c = x == y; break;
                    case QDB_CMP_NE: c = x != y; break;
                    case QDB_CMP_LT: c = x < y; break;
                    case QDB_CMP_LE: c = x <= y; break;
                    case QDB_CMP_GT: c = x > y; break;
                    default: c = x >= y; break;
                    }
                } else {
                    i128 x = a.i, y = b.i;
                    switch (in->ImmLo) {
                    case QDB_CMP_EQ: c = x == y; break;
                    case QDB_CMP_NE: c = x != y; break;
                    case QDB_CMP_LT: c = x < y; break;
                    case QDB_CMP_LE: c = x <= y; break;
                    case QDB_CMP_GT: c = x > y; break;
                    default: c = x >= y; break;
                    }
                }
            }
            out.i = c;
            break;
        }
        case QDB_EXPR_AND: {
            int af = a.valid && a.i == 0, bf = b.valid && b.i == 0;
            out.cls = QDB_REG_BOOL;
            if (af || bf) { out.valid = 1; out.i = 0; }
            else if (!a.valid || !b.valid) { out.valid = 0; out.i = 0; }
            else { out.valid = 1; out.i = 1; }
            break;
        }
        case QDB_EXPR_OR: {
            int at = a.valid && a.i != 0, bt = b.valid && b.i != 0;
            out.cls = QDB_REG_BOOL;
            if (at || bt) { out.valid = 1; out.i = 1; }
            else if (!a.valid || !b.valid) { out.valid = 0; out.i = 0; }
            else { out.valid = 1; out.i = 0; }
            break;
        }
        case QDB_EXPR_NOT:
            out.cls = QDB_REG_BOOL;
            out.valid = a.valid;
            out.i = a.valid ? (a.i == 0) : 0;
            break;
        case QDB_EXPR_IS_NULL:
            out.cls = QDB_REG_BOOL;
            out.valid = 1;
            out.i = !a.valid;
            break;
        case QDB_EXPR_DIV:
            /* IEEE f64 division: AVG = SUM / COUNT (qdb/plan/sql_lowering.cpp:799-814). */
            out.cls = QDB_REG_F64;
            out.valid = a.valid && b.valid;
            if (out.valid) {
                double x = a.cls == QDB_REG_F64 ? a.f : (double)a.i;
                double y = b.cls == QDB_REG_F64 ? b.f : (double)b.i;
                out.f = x / y;
            }
            break;
        case QDB_EXPR_TO_F64:
            out.cls = QDB_REG_F64;
            out.valid = a.valid;
            if (a.cls == QDB_REG_F64) out.f = a.f;
            else if (a.valid) out.f = oracle_decimal_to_f64((int64_t)(uint64_t)(u128)a.i,
                (int64_t)(uint64_t)((u128)a.i >> 64), in->ImmLo);
            break;
        default:
            break;
        }
        *d = out;
    }
}

/* Filter: sel[i] = (input_sel == NULL || input_sel[i]) ? pred IS TRUE : 0. */
void oracle_filter(const QdbFilterSpec* spec, const TRowSet* rs, const uint8_t* input_sel, uint8_t* out_sel)
{
    Reg r[QDB_EXPR_MAX_REGS];
    for (int64_t i = 0; i < rs->RowCount; ++i) {
        if (input_sel && !input_sel[i]) {
            out_sel[i] = 0;
            continue;
        }
        memset(r, 0, sizeof(r));
        eval_program(&spec->Predicate, &spec->Input, rs->Columns, rs->ColumnCount, i, r);
        const Reg* res = &r[spec->Predicate.Result];
        out_sel[i] = (uint8_t)(res->valid && res->i != 0);
    }
}

static void set_bit(uint8_t* bitmap, int64_t i, int v)
{
    if (v) bitmap[i >> 3] |= (uint8_t)(1u << (i & 7));
    else bitmap[i >> 3] &= (uint8_t)~(1u << (i & 7));
}

/* The row-error check of a DECIMAL(p <= 18) store: qdb_decimal64_narrow_checked
 * (qdb/modules/qumirdb.oz:1236-1245) calls qdb_decimal64_overflow(p) when
 * |value| >= 10^p, which raises, or returns 0 while row errors are deferred
 * (qdb/modules/qumirdb_runtime.cpp:662-667). Returns p on failure, else 0. */
static int narrow_fails(const QdbValueType* t, const Reg* v)
{
    if (t->Kind != QDB_KIND_DECIMAL || t->Width != 8 || !v->valid || v->cls == QDB_REG_F64) return 0;
    const int p = t->Precision ? t->Precision : 18;
    const i128 lim = pow10_i128(p);
    return (v->i >= lim || v->i <= -lim) ? p : 0;
}

static void store_value(void* data, const QdbValueType* t, int64_t row, const Reg* v)
{
    if (t->Kind == QDB_KIND_BOOL) {
        set_bit((uint8_t*)data, row, v->valid && v->i != 0);
        return;
    }
    if (t->Kind == QDB_KIND_FLOAT) {
        double f = v->cls == QDB_REG_F64 ? v->f : (double)v->i;
        if (t->Width == 4) ((float*)data)[row] = (float)f;
        else ((double*)data)[row] = f;
        return;
    }
    i128 x = v->valid ? v->i : 0;
    switch (t->Width) {
    case 1: ((int8_t*)data)[row] = (int8_t)x; break;
    case 2: ((int16_t*)data)[row] = (int16_t)x; break;
    case 4: ((int32_t*)data)[row] = (int32_t)x; break;
    case 8: ((int64_t*)data)[row] = (int64_t)x; break;
    default: memcpy((char*)data + 16 * row, &x, 16); break;
    }
}

/* Project: out[k] data, out[NumOutputs + k] validity of nullable outputs.
 * Returns 0, or the precision of the first row error (the failing value is
 * stored as 0, as under deferral). */
int oracle_project(const QdbProjectSpec* spec, const TRowSet* in, void** out, const uint8_t* input_sel)
{
    Reg r[QDB_EXPR_MAX_REGS];
    int failed = 0;
    for (int64_t i = 0; i < in->RowCount; ++i) {
        if (input_sel && !input_sel[i]) continue;
        for (int k = 0; k < spec->NumOutputs; ++k) {
            memset(r, 0, sizeof(r));
            eval_program(&spec->Outputs[k], &spec->Input, in->Columns, in->ColumnCount, i, r);
            Reg* v = &r[spec->Outputs[k].Result];
            const int p = narrow_fails(&spec->OutputTypes[k], v);
            if (p) {
                if (!failed) failed = p;
                v->i = 0;
            }
            store_value(out[k], &spec->OutputTypes[k], i, v);
            if (spec->OutputTypes[k].Nullable) {
                set_bit((uint8_t*)out[spec->NumOutputs + k], i, v->valid);
            }
        }
    }
    return failed;
}

/* ------------------------------------------------------------------------ */
/* Key hashing, HashKeyValue (qdb/kernel/gen.cpp:233-379)                    */
/* ------------------------------------------------------------------------ */

uint64_t oracle_mix_u64(uint64_t h)
{
    h ^= h >> 12;
    h ^= h << 25;
    h ^= h >> 27;
    return h * 2685821657736338717ull;
}

static uint64_t canonical_f64_bits(double v)
{
    uint64_t b;
    memcpy(&b, &v, 8);
    if ((b & 0x7fffffffffffffffull) == 0) return 0;
    if ((b & 0x7ff0000000000000ull) == 0x7ff0000000000000ull && (b & 0x000fffffffffffffull) != 0) {
        return 0x7ff8000000000000ull;
    }
    return b;
}

/* The leaf hash of one key field (valid only). */
static uint64_t leaf_hash(const TColumn* c, const QdbValueType* t, int64_t row)
{
    if (t->Kind == QDB_KIND_FLOAT) {
        if (t->Width == 4) {
            float f = ((const float*)c->Data)[row];
            uint32_t b;
            memcpy(&b, &f, 4);
            if ((b & 0x7fffffffu) == 0) b = 0;
            else if ((b & 0x7f800000u) == 0x7f800000u && (b & 0x007fffffu) != 0) b = 0x7fc00000u;
            return oracle_mix_u64((uint64_t)b);
        }
        return oracle_mix_u64(canonical_f64_bits(((const double*)c->Data)[row]));
    }
    if (t->Kind == QDB_KIND_BOOL) {
        return (uint64_t)bit_at((const uint8_t*)c->Data, row, c->DataBitOffset) * 2685821657736338717ull;
    }
    if (t->Width == 16) {
        i128 v = col_int(c, t, row);
        uint64_t lo = oracle_mix_u64((uint64_t)(u128)v);
        uint64_t hi = oracle_mix_u64((uint64_t)((u128)v >> 64));
        return lo ^ (hi + 0x9E3779B97F4A7C15ull + (lo << 6) + (lo >> 2));
    }
    /* Zero-extension of the value's own width (a cast to the unsigned type). */
    uint64_t bits = 0;
    switch (t->Width) {
    case 1: bits = ((const uint8_t*)c->Data)[row]; break;
    case 2: bits = ((const uint16_t*)c->Data)[row]; break;
    case 4: bits = ((const uint32_t*)c->Data)[row]; break;
    default: bits = ((const uint64_t*)c->Data)[row]; break;
    }
    return oracle_mix_u64(bits);
}

/* rh_hash of a key (qdb/kernel/gen.cpp:344-373 over the key struct built by
 * BuildAggregateKeyDescriptor, qdb/kernel/aggregate_key.cpp:246-272). The key
 * is always a struct, even with one field: for each key column a nullable
 * column first contributes its `valid_i` bool field (hashed as a bool leaf,
 * valid * 2685821657736338717, gen.cpp:319-325), then its `key_i` field
 * (the leaf hash, or 0 when the row's value is NULL, gen.cpp:357-364); each
 * field hash enters the boost-style ordered combine
 * h ^= fh + 0x9E3779B97F4A7C15 + (h << 6) + (h >> 2), starting from h = 0. */
uint64_t oracle_key_hash(const TColumn* cols, const QdbValueType* types, const int32_t* key_cols, int nkeys, int64_t row)
{
    uint64_t h = 0;
    for (int k = 0; k < nkeys; ++k) {
        const TColumn* c = &cols[key_cols[k]];
        const QdbValueType* t = &types[key_cols[k]];
        int valid = 1;
        if (t->Nullable) {
            valid = col_valid(c, row);
            const uint64_t vh = valid ? 2685821657736338717ull : 0ull;
            h ^= vh + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
        }
        const uint64_t fh = valid ? leaf_hash(c, t, row) : 0;
        h ^= fh + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
    }
    return h;
}

/* ------------------------------------------------------------------------ */
/* Exact and compensated f64 summation                                       */
/* ------------------------------------------------------------------------ */

/* A fixed-point accumulator over every finite double: bit i of the sum has
 * weight 2^(i - 1074 - 64). 32-bit chunks in signed 64-bit limbs absorb up
 * to 2^31 additions without carry propagation. */
#define SUPER_LIMBS 72
typedef struct {
    int64_t limb[SUPER_LIMBS];
    int64_t adds;
    int nan, pinf, ninf;
} Super;

static void super_normalize(Super* s)
{
    for (int i = 0; i < SUPER_LIMBS - 1; ++i) {
        int64_t carry = s->limb[i] >> 32; /* arithmetic shift: floor division */
        s->limb[i] -= carry * ((int64_t)1 << 32);
        s->limb[i + 1] += carry;
    }
    s->adds = 0;
}

static void super_add(Super* s, double x)
{
    uint64_t b;
    memcpy(&b, &x, 8);
    int exp = (int)((b >> 52) & 0x7ff);
    uint64_t man = b & 0x000fffffffffffffull;
    int neg = (int)(b >> 63);
    if (exp == 0x7ff) {
        if (man) s->nan = 1;
        else if (neg) s->ninf = 1;
        else s->pinf = 1;
        return;
    }
    if (exp == 0) {
        exp = 1;
    } else {
        man |= 1ull << 52;
    }
    if (man == 0) return;
    /* value = man * 2^(exp - 1075); bit position in the accumulator: */
    int pos = exp - 1075 + 1074 + 64; /* >= 64 */
    int li = pos / 32, sh = pos % 32;
    u128 m = (u128)man << sh; /* at most 85 bits */
    for (int j = 0; j < 3 && li + j < SUPER_LIMBS; ++j) {
        int64_t part = (int64_t)(uint64_t)(m & 0xffffffffu);
        s->limb[li + j] += neg ? -part : part;
        m >>= 32;
    }
    if (++s->adds >= ((int64_t)1 << 30)) super_normalize(s);
}

static double super_round(Super* s)
{
    if (s->nan || (s->pinf && s->ninf)) return NAN;
    if (s->pinf) return INFINITY;
    if (s->ninf) return -INFINITY;
    super_normalize(s);
    /* After normalisation every limb but the top one is in [0, 2^32) and the
     * top limb carries the sign. A negative sum is negated limb-wise and
     * normalised again, which leaves its magnitude. */
    int neg = s->limb[SUPER_LIMBS - 1] < 0;
    if (neg) {
        for (int i = 0; i < SUPER_LIMBS; ++i) s->limb[i] = -s->limb[i];
        super_normalize(s);
    }
    uint32_t mag[SUPER_LIMBS];
    for (int i = 0; i < SUPER_LIMBS; ++i) mag[i] = (uint32_t)s->limb[i];
    /* Highest set bit. */
    int top = -1;
    for (int i = SUPER_LIMBS - 1; i >= 0 && top < 0; --i) {
        if (mag[i]) top = i * 32 + 31 - __builtin_clz(mag[i]);
    }
    if (top < 0) return 0.0;
    /* Result LSB position: 53 significant bits, but not below the subnormal
     * limit (accumulator bit 64 is 2^-1074). */
    int lsb = top - 52;
    if (lsb < 64) lsb = 64;
#define MAGBIT(p) (((p) < 0) ? 0 : (int)((mag[(p) / 32] >> ((p) % 32)) & 1u))
    uint64_t q = 0;
    for (int p = top; p >= lsb; --p) q = (q << 1) | (uint64_t)MAGBIT(p);
    int guard = lsb - 1 >= 0 ? MAGBIT(lsb - 1) : 0;
    int sticky = 0;
    for (int p = lsb - 2; p >= 0 && !sticky; --p) sticky = MAGBIT(p);
#undef MAGBIT
    if (guard && (sticky || (q & 1))) q += 1;
    double r = ldexp((double)q, lsb - 64 - 1074);
    return neg ? -r : r;
}

/* Exact sum of n doubles, rounded once to nearest-even. */
double oracle_exact_sum(const double* v, int64_t n)
{
    Super* s = (Super*)calloc(1, sizeof(Super));
    for (int64_t i = 0; i < n; ++i) super_add(s, v[i]);
    double r = super_round(s);
    free(s);
    return r;
}

/* Kahan-Babuska-Neumaier compensated sum, in order. */
double oracle_kahan_sum(const double* v, int64_t n)
{
    double sum = 0.0, c = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        double t = sum + v[i];
        if (fabs(sum) >= fabs(v[i])) c += (sum - t) + v[i];
        else c += (v[i] - t) + sum;
        sum = t;
    }
    return sum + c;
}

/* ------------------------------------------------------------------------ */
/* Aggregate: swiss table + reducers                                         */
/* ------------------------------------------------------------------------ */

typedef struct {
    uint8_t valid;
    i128 v; /* integer / decimal value, or canonical float bits */
} KeyField;

typedef struct {
    KeyField f[QDB_AGG_MAX_KEYS];
} GroupKey;

enum { R_COUNT, R_SUM_I64, R_SUM_I128, R_SUM_F64, R_MINMAX_INT, R_MINMAX_F64 };

typedef struct {
    int kind;      /* R_* */
    int func;      /* QDB_AGG_* */
    int column;
    int nullable;  /* argument nullable: keep a valid counter */
} Reducer;

typedef struct {
    int64_t rows;          /* rows that reached the group */
    int64_t valid;         /* valid argument values (nullable args) */
    i128 i;                /* count / integer sum / i128 sum / min-max int */
    double f;              /* arrival-order f64 sum / min-max */
    double kc;             /* Kahan-Neumaier compensation */
} AggState;

typedef struct OracleAgg {
    QdbAggregateSpec spec;
    Reducer red[QDB_AGG_MAX_AGGS];
    int fsum_mode;         /* 0 arrival order, 1 Kahan-Neumaier, 2 exact */
    /* swiss table */
    uint8_t* ctrl;
    int64_t* slot_dense;   /* physical slot -> dense id */
    int64_t capacity;
    /* dense groups */
    GroupKey* keys;
    AggState* states;      /* [dense * num_aggs + a] */
    int64_t size, dense_cap;
    /* exact mode: (dense id, value) of every f64 SUM contribution per agg */
    int64_t* ex_gid[QDB_AGG_MAX_AGGS];
    double* ex_val[QDB_AGG_MAX_AGGS];
    int64_t ex_n[QDB_AGG_MAX_AGGS], ex_cap[QDB_AGG_MAX_AGGS];
    uint64_t* hashes;      /* hash per dense group (rehash) */
} OracleAgg;

static int key_equal(const GroupKey* a, const GroupKey* b, int nkeys)
{
    for (int k = 0; k < nkeys; ++k) {
        if (a->f[k].valid != b->f[k].valid) return 0;
        if (a->f[k].valid && a->f[k].v != b->f[k].v) return 0;
    }
    return 1;
}

static void read_key(const OracleAgg* ag, const TRowSet* rs, int64_t row, GroupKey* key)
{
    memset(key, 0, sizeof(*key));
    for (int k = 0; k < ag->spec.NumKeys; ++k) {
        int c = ag->spec.KeyColumns[k];
        const TColumn* col = &rs->Columns[c];
        const QdbValueType* t = &ag->spec.Input.Types[c];
        key->f[k].valid = (uint8_t)col_valid(col, row);
        if (!key->f[k].valid) continue;
        if (t->Kind == QDB_KIND_FLOAT) {
            double d = col_f64(col, t, row);
            key->f[k].v = (i128)canonical_f64_bits(d);
        } else {
            key->f[k].v = col_int(col, t, row);
        }
    }
}

/* SWAR group match (swiss_group.oz:27-40): bytes of an 8-slot group equal
 * to h2, as 0x80 per matching byte. */
static uint64_t swar_match(const uint8_t* g, uint64_t h2)
{
    uint64_t word;
    memcpy(&word, g, 8);
    const uint64_t lsbs = 0x0101010101010101ull, msbs = lsbs << 7;
    uint64_t x = word ^ (lsbs * h2);
    return (x - lsbs) & (msbs ^ (x & msbs));
}

static uint64_t swar_empty(const uint8_t* g)
{
    uint64_t word;
    memcpy(&word, g, 8);
    return word & (0x0101010101010101ull << 7);
}

static int64_t swiss_find(OracleAgg* ag, const GroupKey* key, uint64_t hash, int64_t* empty_slot)
{
    const int64_t num_groups = ag->capacity >> 3, group_mask = num_groups - 1;
    const uint64_t h2 = hash & 0x7f;
    int64_t g = (int64_t)(hash >> 7) & group_mask, step = 0;
    *empty_slot = -1;
    for (int64_t probes = 0; probes < num_groups; ++probes) {
        const uint8_t* grp = ag->ctrl + (g << 3);
        uint64_t m = swar_match(grp, h2);
        while (m) {
            int64_t slot = (g << 3) + (__builtin_ctzll(m) >> 3);
            int64_t d = ag->slot_dense[slot];
            if (key_equal(&ag->keys[d], key, ag->spec.NumKeys)) return d;
            m &= m - 1;
        }
        uint64_t e = swar_empty(grp);
        if (e) {
            *empty_slot = (g << 3) + (__builtin_ctzll(e) >> 3);
            return -1;
        }
        step += 1;
        g = (g + step) & group_mask;
    }
    return -1;
}

static void swiss_alloc(OracleAgg* ag, int64_t capacity)
{
    ag->capacity = capacity;
    ag->ctrl = (uint8_t*)malloc((size_t)capacity);
    memset(ag->ctrl, 0x80, (size_t)capacity);
    ag->slot_dense = (int64_t*)malloc((size_t)capacity * 8);
}

static void swiss_rehash(OracleAgg* ag, int64_t capacity)
{
    free(ag->ctrl);
    free(ag->slot_dense);
    swiss_alloc(ag, capacity);
    for (int64_t d = 0; d < ag->size; ++d) {
        int64_t empty;
        swiss_find(ag, &ag->keys[d], ag->hashes[d], &empty);
        ag->ctrl[empty] = (uint8_t)(ag->hashes[d] & 0x7f);
        ag->slot_dense[empty] = d;
    }
}

static int64_t swiss_upsert(OracleAgg* ag, const GroupKey* key, uint64_t hash)
{
    int64_t empty;
    int64_t d = swiss_find(ag, key, hash, &empty);
    if (d >= 0) return d;
    /* aht_insert_grows: grow before crossing 7/8 of the capacity. */
    if (ag->size + 1 > ag->capacity - (ag->capacity >> 3)) {
        swiss_rehash(ag, ag->capacity * 2);
        swiss_find(ag, key, hash, &empty);
    }
    if (ag->size == ag->dense_cap) {
        ag->dense_cap = ag->dense_cap ? ag->dense_cap * 2 : 64;
        ag->keys = (GroupKey*)realloc(ag->keys, (size_t)ag->dense_cap * sizeof(GroupKey));
        ag->hashes = (uint64_t*)realloc(ag->hashes, (size_t)ag->dense_cap * 8);
        ag->states = (AggState*)realloc(ag->states,
            (size_t)ag->dense_cap * (size_t)(ag->spec.NumAggs ? ag->spec.NumAggs : 1) * sizeof(AggState));
    }
    d = ag->size++;
    ag->keys[d] = *key;
    ag->hashes[d] = hash;
    memset(&ag->states[d * ag->spec.NumAggs], 0, (size_t)ag->spec.NumAggs * sizeof(AggState));
    ag->ctrl[empty] = (uint8_t)(hash & 0x7f);
    ag->slot_dense[empty] = d;
    return d;
}

static int reducer_of(const QdbAggregateSpec* spec, int a, Reducer* r)
{
    const QdbAggregateFn* fn = &spec->Aggs[a];
    r->func = fn->Func;
    r->column = fn->Column;
    r->nullable = 0;
    if (fn->Func == QDB_AGG_COUNT_STAR) {
        r->kind = R_COUNT;
        return 0;
    }
    const QdbValueType* t = &spec->Input.Types[fn->Column];
    r->nullable = t->Nullable;
    if (fn->Func == QDB_AGG_COUNT) {
        r->kind = R_COUNT;
        return 0;
    }
    if (t->Kind == QDB_KIND_FLOAT) {
        r->kind = fn->Func == QDB_AGG_SUM ? R_SUM_F64 : R_MINMAX_F64;
        return 0;
    }
    if (fn->Func == QDB_AGG_SUM) {
        r->kind = t->Kind == QDB_KIND_DECIMAL ? R_SUM_I128 : R_SUM_I64;
        if (t->Kind == QDB_KIND_DATE) return -1;
        return 0;
    }
    r->kind = R_MINMAX_INT;
    return 0;
}

OracleAgg* oracle_agg_new(const QdbAggregateSpec* spec, int64_t capacity, int fsum_mode)
{
    OracleAgg* ag = (OracleAgg*)calloc(1, sizeof(OracleAgg));
    ag->spec = *spec;
    ag->fsum_mode = fsum_mode;
    /* Computed columns follow the input columns. */
    for (int k = 0; k < spec->NumComputed && spec->Input.NumColumns + k < 64; ++k) {
        ag->spec.Input.Types[spec->Input.NumColumns + k] = spec->ComputedTypes[k];
    }
    for (int a = 0; a < spec->NumAggs; ++a) {
        if (reducer_of(&ag->spec, a, &ag->red[a]) != 0) {
            free(ag);
            return NULL;
        }
    }
    if (capacity < 8) capacity = 8;
    int64_t c = 8;
    while (c < capacity) c *= 2;
    swiss_alloc(ag, c);
    /* Empty-input groups (qdb/kernel/gen.cpp:2218-2269): the leading key
     * holds the id, the other fields their zero value (NULL when nullable);
     * upserted with rh_hash of that key, reducers untouched. */
    for (int g = 0; g < spec->NumEmptyInputGroups; ++g) {
        GroupKey key;
        memset(&key, 0, sizeof(key));
        int64_t vals[QDB_AGG_MAX_KEYS][2];
        uint8_t masks[QDB_AGG_MAX_KEYS][8];
        TColumn cols[64];
        memset(vals, 0, sizeof(vals));
        memset(masks, 0, sizeof(masks));
        memset(cols, 0, sizeof(cols));
        for (int k = 0; k < spec->NumKeys; ++k) {
            const int kc = spec->KeyColumns[k];
            const int nullable = ag->spec.Input.Types[kc].Nullable;
            key.f[k].valid = (uint8_t)(k == 0 || !nullable);
            key.f[k].v = k == 0 ? (i128)spec->EmptyInputGroups[g] : 0;
            /* One-row columns of the seed key, for its hash. */
            const int64_t v = k == 0 ? spec->EmptyInputGroups[g] : 0;
            const int w = ag->spec.Input.Types[kc].Width;
            memcpy(vals[k], &v, w < 8 ? (size_t)w : 8);
            cols[kc].Data = (char*)vals[k];
            if (nullable) {
                masks[k][0] = key.f[k].valid;
                cols[kc].Mask = masks[k];
            }
        }
        const uint64_t h = oracle_key_hash(cols, ag->spec.Input.Types, spec->KeyColumns, spec->NumKeys, 0);
        (void)swiss_upsert(ag, &key, h);
    }
    return ag;
}

static void ex_push(OracleAgg* ag, int a, int64_t gid, double v)
{
    if (ag->ex_n[a] == ag->ex_cap[a]) {
        ag->ex_cap[a] = ag->ex_cap[a] ? ag->ex_cap[a] * 2 : 1024;
        ag->ex_gid[a] = (int64_t*)realloc(ag->ex_gid[a], (size_t)ag->ex_cap[a] * 8);
        ag->ex_val[a] = (double*)realloc(ag->ex_val[a], (size_t)ag->ex_cap[a] * 8);
    }
    ag->ex_gid[a][ag->ex_n[a]] = gid;
    ag->ex_val[a][ag->ex_n[a]] = v;
    ag->ex_n[a]++;
}

static void apply_row(OracleAgg* ag, const TRowSet* rs, int64_t row, int64_t d)
{
    for (int a = 0; a < ag->spec.NumAggs; ++a) {
        const Reducer* r = &ag->red[a];
        AggState* s = &ag->states[d * ag->spec.NumAggs + a];
        int is_new = s->rows == 0;
        s->rows += 1;
        if (r->func == QDB_AGG_COUNT_STAR) {
            s->i += 1;
            continue;
        }
        const TColumn* c = &rs->Columns[r->column];
        const QdbValueType* t = &ag->spec.Input.Types[r->column];
        int valid = col_valid(c, row);
        if (r->func == QDB_AGG_COUNT) {
            s->i += valid;
            continue;
        }
        if (!valid) continue;
        /* With a nullable argument the first valid value seeds the state
         * (gen.cpp:3533-3580); otherwise the group's first row does. */
        int seed = r->nullable ? s->valid == 0 : is_new;
        s->valid += 1;
        switch (r->kind) {
        case R_SUM_I64:
            s->i = (i128)(int64_t)((uint64_t)(int64_t)s->i + (uint64_t)(int64_t)col_int(c, t, row));
            break;
        case R_SUM_I128:
            s->i = (i128)((u128)s->i + (u128)col_int(c, t, row));
            break;
        case R_SUM_F64: {
            double v = col_f64(c, t, row);
            if (ag->fsum_mode == 1) {
                double tt = s->f + v;
                if (fabs(s->f) >= fabs(v)) s->kc += (s->f - tt) + v;
                else s->kc += (v - tt) + s->f;
                s->f = tt;
            } else if (ag->fsum_mode == 2) {
                ex_push(ag, a, d, v);
            } else {
                s->f = s->f + v;
            }
            break;
        }
        case R_MINMAX_INT: {
            i128 v = col_int(c, t, row);
            if (seed) s->i = v;
            else if (r->func == QDB_AGG_MIN ? v < s->i : v > s->i) s->i = v;
            break;
        }
        case R_MINMAX_F64: {
            double v = col_f64(c, t, row);
            if (seed) s->f = v;
            else if (r->func == QDB_AGG_MIN ? v < s->f : v > s->f) s->f = v;
            break;
        }
        default:
            break;
        }
    }
}

/* agg_dispatch(Update) over one batch: selected rows only; the batch's
 * precomputed hashes are used when present. */
static void agg_update_rows(OracleAgg* ag, const TRowSet* rs)
{
    GroupKey key;
    for (int64_t i = 0; i < rs->RowCount; ++i) {
        if (rs->Selection && !rs->Selection[i]) continue;
        read_key(ag, rs, i, &key);
        uint64_t h = ag->spec.NumKeys == 0 ? 0
            : rs->Hash ? rs->Hash[i]
            : oracle_key_hash(rs->Columns, ag->spec.Input.Types, ag->spec.KeyColumns, ag->spec.NumKeys, i);
        int64_t d = swiss_upsert(ag, &key, h);
        apply_row(ag, rs, i, d);
    }
}

/* A spec with a fused predicate and computed columns is the composition of
 * the filter kernel (selection), the project kernel (computed columns after
 * the input columns) and the aggregate over the projected batch. */
void oracle_agg_update(OracleAgg* ag, const TRowSet* rs)
{
    const QdbAggregateSpec* spec = &ag->spec;
    if (!spec->HasPredicate && spec->NumComputed == 0) {
        agg_update_rows(ag, rs);
        return;
    }
    const int nin = (int)rs->ColumnCount;
    const int total = nin + spec->NumComputed;
    const int64_t n = rs->RowCount;
    TColumn* cols = (TColumn*)calloc((size_t)total, sizeof(TColumn));
    memcpy(cols, rs->Columns, (size_t)nin * sizeof(TColumn));
    void* owned[2 * QDB_PROJECT_MAX_OUTPUTS + 1];
    int nowned = 0;
    QdbProjectSpec ps;
    memset(&ps, 0, sizeof(ps));
    ps.Input = spec->Input;
    ps.Input.NumColumns = nin;
    ps.NumOutputs = spec->NumComputed;
    void* outs[2 * QDB_PROJECT_MAX_OUTPUTS];
    memset(outs, 0, sizeof(outs));
    for (int k = 0; k < spec->NumComputed; ++k) {
        const QdbValueType* t = &spec->ComputedTypes[k];
        ps.OutputTypes[k] = *t;
        ps.Outputs[k] = spec->Computed[k];
        size_t bytes = t->Kind == QDB_KIND_BOOL ? (size_t)(n + 7) / 8 + 8 : (size_t)n * (size_t)t->Width + 16;
        outs[k] = calloc(bytes, 1);
        owned[nowned++] = outs[k];
        cols[nin + k].Data = (char*)outs[k];
        if (t->Nullable) {
            outs[spec->NumComputed + k] = calloc((size_t)(n + 7) / 8 + 8, 1);
            owned[nowned++] = outs[spec->NumComputed + k];
            cols[nin + k].Mask = (const uint8_t*)outs[spec->NumComputed + k];
        }
    }
    if (spec->NumComputed > 0) oracle_project(&ps, rs, outs, NULL);
    uint8_t* sel = NULL;
    if (spec->HasPredicate) {
        QdbFilterSpec fs;
        memset(&fs, 0, sizeof(fs));
        fs.Input = spec->Input;
        fs.Input.NumColumns = nin;
        fs.Predicate = spec->Predicate;
        sel = (uint8_t*)calloc((size_t)(n ? n : 1), 1);
        oracle_filter(&fs, rs, rs->Selection, sel);
        owned[nowned++] = sel;
    }
    TRowSet ext = *rs;
    ext.Columns = cols;
    ext.ColumnCount = total;
    if (sel) ext.Selection = sel;
    agg_update_rows(ag, &ext);
    for (int i = 0; i < nowned; ++i) free(owned[i]);
    free(cols);
}

int64_t oracle_agg_size(const OracleAgg* ag)
{
    return ag->size;
}

static int cmp_gid(const void* x, const void* y)
{
    const int64_t* a = (const int64_t*)x;
    const int64_t* b = (const int64_t*)y;
    return a[0] < b[0] ? -1 : a[0] > b[0] ? 1 : (a[1] < b[1] ? -1 : a[1] > b[1]);
}

/* Result types, AggregateResultType (qdb/plan/ops/aggregate.cpp:21-61). */
static const QdbValueType* spec_col_type(const QdbAggregateSpec* spec, int c)
{
    if (c >= spec->Input.NumColumns && c - spec->Input.NumColumns < spec->NumComputed) {
        return &spec->ComputedTypes[c - spec->Input.NumColumns];
    }
    return &spec->Input.Types[c];
}

void oracle_agg_output_types(const QdbAggregateSpec* spec, QdbValueType* out)
{
    int n = 0;
    for (int k = 0; k < spec->NumKeys; ++k) out[n++] = *spec_col_type(spec, spec->KeyColumns[k]);
    for (int a = 0; a < spec->NumAggs; ++a) {
        QdbValueType t;
        memset(&t, 0, sizeof(t));
        const QdbAggregateFn* fn = &spec->Aggs[a];
        if (fn->Func == QDB_AGG_COUNT_STAR || fn->Func == QDB_AGG_COUNT) {
            t.Kind = QDB_KIND_SIGNED;
            t.Width = 8;
        } else {
            const QdbValueType* in = spec_col_type(spec, fn->Column);
            t.Nullable = in->Nullable || spec->NumEmptyInputGroups > 0;
            if (in->Kind == QDB_KIND_FLOAT) {
                t.Kind = QDB_KIND_FLOAT;
                t.Width = 8;
            } else if (in->Kind == QDB_KIND_DECIMAL) {
                t.Kind = QDB_KIND_DECIMAL;
                t.Scale = in->Scale;
                t.Width = fn->Func == QDB_AGG_SUM ? 16 : in->Width;
            } else if (in->Kind == QDB_KIND_DATE) {
                t.Kind = QDB_KIND_DATE;
                t.Width = 8;
            } else {
                t.Kind = QDB_KIND_SIGNED;
                t.Width = 8;
            }
        }
        out[n++] = t;
    }
}

static void destroy_host_rowset(TRowSet* rs)
{
    void** owned = (void**)rs->Private;
    if (owned) {
        int64_t n = (int64_t)(intptr_t)owned[0];
        for (int64_t i = 1; i <= n; ++i) free(owned[i]);
        free(owned);
    }
    free(rs->Columns);
    rs->Columns = NULL;
    rs->Private = NULL;
}

/* agg_finish_rowset: groups in dense (first-occurrence) order; keys then one
 * column per aggregate. Returns the group count. Frees the table. */
int64_t oracle_agg_finish(OracleAgg* ag, TRowSet* out)
{
    QdbAggregateSpec full = ag->spec;
    const QdbAggregateSpec* spec = &full;
    const int64_t n = ag->size;
    const int ncols = spec->NumKeys + spec->NumAggs;
    QdbValueType types[QDB_AGG_MAX_KEYS + QDB_AGG_MAX_AGGS];
    oracle_agg_output_types(spec, types);
    double* exact[QDB_AGG_MAX_AGGS] = {0};
    for (int a = 0; a < spec->NumAggs; ++a) {
        if (ag->red[a].kind != R_SUM_F64 || ag->fsum_mode != 2) continue;
        int64_t m = ag->ex_n[a];
        int64_t* pairs = (int64_t*)malloc((size_t)(m ? m : 1) * 16);
        for (int64_t i = 0; i < m; ++i) {
            pairs[2 * i] = ag->ex_gid[a][i];
            pairs[2 * i + 1] = i;
        }
        qsort(pairs, (size_t)m, 16, cmp_gid);
        exact[a] = (double*)calloc((size_t)(n ? n : 1), 8);
        double* buf = (double*)malloc((size_t)(m ? m : 1) * 8);
        for (int64_t i = 0; i < m;) {
            int64_t j = i, g = pairs[2 * i];
            while (j < m && pairs[2 * j] == g) {
                buf[j - i] = ag->ex_val[a][pairs[2 * j + 1]];
                ++j;
            }
            exact[a][g] = oracle_exact_sum(buf, j - i);
            i = j;
        }
        free(buf);
        free(pairs);
    }
    out->Columns = (TColumn*)calloc((size_t)(ncols ? ncols : 1), sizeof(TColumn));
    out->ColumnCount = ncols;
    out->RowCount = n;
    out->Selection = NULL;
    out->Hash = NULL;
    out->RefCount = 1;
    out->Destroy = destroy_host_rowset;
    void** owned = (void**)calloc((size_t)(2 * ncols + 1), sizeof(void*));
    int64_t nowned = 0;
    const size_t mask_bytes = (size_t)((n + 7) / 8) + 1;
    for (int c = 0; c < ncols; ++c) {
        const QdbValueType* t = &types[c];
        size_t width = (size_t)(t->Kind == QDB_KIND_BOOL ? 1 : t->Width);
        char* data = (char*)calloc((size_t)(n ? n : 1), width ? width : 1);
        if (t->Kind == QDB_KIND_BOOL) {
            free(data);
            data = (char*)calloc(mask_bytes, 1);
        }
        uint8_t* mask = NULL;
        if (t->Nullable) mask = (uint8_t*)calloc(mask_bytes, 1);
        owned[++nowned] = data;
        if (mask) owned[++nowned] = mask;
        out->Columns[c].Data = data;
        out->Columns[c].Mask = mask;
        for (int64_t d = 0; d < n; ++d) {
            Reg v;
            memset(&v, 0, sizeof(v));
            v.valid = 1;
            if (c < spec->NumKeys) {
                const KeyField* f = &ag->keys[d].f[c];
                v.valid = f->valid;
                if (t->Kind == QDB_KIND_FLOAT) {
                    uint64_t bits = (uint64_t)f->v;
                    memcpy(&v.f, &bits, 8);
                    v.cls = QDB_REG_F64;
                } else {
                    v.i = f->v;
                    v.cls = QDB_REG_INT;
                }
            } else {
                int a = c - spec->NumKeys;
                const Reducer* r = &ag->red[a];
                const AggState* s = &ag->states[d * spec->NumAggs + a];
                if (r->kind == R_COUNT) {
                    v.i = s->i;
                } else {
                    int64_t contributions = r->nullable ? s->valid : s->rows;
                    v.valid = contributions > 0;
                    if (r->kind == R_SUM_F64) {
                        v.cls = QDB_REG_F64;
                        v.f = ag->fsum_mode == 1 ? s->f + s->kc : ag->fsum_mode == 2 ? exact[a][d] : s->f;
                    } else if (r->kind == R_MINMAX_F64) {
                        v.cls = QDB_REG_F64;
                        v.f = s->f;
                    } else {
                        v.i = s->i;
                    }
                }
            }
            store_value(data, t, d, &v);
            if (mask) set_bit(mask, d, v.valid);
        }
    }
    owned[0] = (void*)(intptr_t)nowned;
    out->Private = owned;
    for (int a = 0; a < QDB_AGG_MAX_AGGS; ++a) free(exact[a]);
    return n;
}

void oracle_agg_free(OracleAgg* ag)
{
    if (!ag) return;
    for (int a = 0; a < QDB_AGG_MAX_AGGS; ++a) {
        free(ag->ex_gid[a]);
        free(ag->ex_val[a]);
    }
    free(ag->ctrl);
    free(ag->slot_dense);
    free(ag->keys);
    free(ag->hashes);
    free(ag->states);
    free(ag);
}

void oracle_rowset_release(TRowSet* rs)
{
    if (rs->Destroy) rs->Destroy(rs);
}

/* ------------------------------------------------------------------------ */
/* Row join                                                                  */
/* ------------------------------------------------------------------------ */

typedef struct OracleJoin {
    QdbRowJoinSpec spec;
    /* arena rows in append order */
    uint64_t* hash;     /* first word: the key hash, later the chain link */
    GroupKey* key;
    int64_t* ordinal;   /* build row ordinal over every appended batch */
    int64_t* next;      /* chain of equal keys (index + 1, 0 ends) */
    int64_t rows, cap;
    int64_t appended;   /* rows of the Append batches so far */
    /* entry table: salt << 48 | (arena index + 1) */
    uint64_t* entries;
    int64_t capacity;
    uint64_t mask;
    int shift;
    int64_t keys;
    /* pairs of the last probe: 2 i64 each */
    int64_t* pairs;
    int64_t npairs, pair_cap;
} OracleJoin;

OracleJoin* oracle_join_new(const QdbRowJoinSpec* spec)
{
    OracleJoin* j = (OracleJoin*)calloc(1, sizeof(OracleJoin));
    j->spec = *spec;
    return j;
}

void oracle_join_free(OracleJoin* j)
{
    if (!j) return;
    free(j->hash);
    free(j->key);
    free(j->ordinal);
    free(j->next);
    free(j->entries);
    free(j->pairs);
    free(j);
}

static void join_read_key(const QdbColumnSchema* schema, const int32_t* key_cols, int nkeys,
    const TRowSet* rs, int64_t row, GroupKey* key, int* all_valid)
{
    memset(key, 0, sizeof(*key));
    *all_valid = 1;
    for (int k = 0; k < nkeys; ++k) {
        const TColumn* c = &rs->Columns[key_cols[k]];
        const QdbValueType* t = &schema->Types[key_cols[k]];
        key->f[k].valid = (uint8_t)col_valid(c, row);
        if (!key->f[k].valid) {
            *all_valid = 0;
            continue;
        }
        key->f[k].v = t->Kind == QDB_KIND_FLOAT ? (i128)canonical_f64_bits(col_f64(c, t, row)) : col_int(c, t, row);
    }
}

static uint64_t join_hash(const QdbColumnSchema* schema, const int32_t* key_cols, int nkeys,
    const TRowSet* rs, int64_t row)
{
    if (rs->Hash) return rs->Hash[row];
    if (nkeys == 1) return leaf_hash(&rs->Columns[key_cols[0]], &schema->Types[key_cols[0]], row);
    uint64_t h = 0;
    for (int k = 0; k < nkeys; ++k) {
        uint64_t fh = leaf_hash(&rs->Columns[key_cols[k]], &schema->Types[key_cols[k]], row);
        h ^= fh + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
    }
    return h;
}

/* jr_append_batch (qdb/kernel/join_gen.cpp:2063-2159): every selected row
 * with a valid key becomes an arena row carrying its hash and key. */
void oracle_join_append(OracleJoin* j, const TRowSet* rs)
{
    const QdbRowJoinSpec* s = &j->spec;
    for (int64_t i = 0; i < rs->RowCount; ++i) {
        if (rs->Selection && !rs->Selection[i]) continue;
        GroupKey key;
        int valid;
        join_read_key(&s->Build, s->BuildKeyColumns, s->NumKeys, rs, i, &key, &valid);
        if (!valid) continue;
        if (j->rows == j->cap) {
            j->cap = j->cap ? j->cap * 2 : 1024;
            j->hash = (uint64_t*)realloc(j->hash, (size_t)j->cap * 8);
            j->key = (GroupKey*)realloc(j->key, (size_t)j->cap * sizeof(GroupKey));
            j->ordinal = (int64_t*)realloc(j->ordinal, (size_t)j->cap * 8);
            j->next = (int64_t*)realloc(j->next, (size_t)j->cap * 8);
        }
        j->hash[j->rows] = join_hash(&s->Build, s->BuildKeyColumns, s->NumKeys, rs, i);
        j->key[j->rows] = key;
        j->ordinal[j->rows] = j->appended + i;
        j->next[j->rows] = 0;
        j->rows++;
    }
    j->appended += rs->RowCount;
}

static uint64_t rt_mix(uint64_t h) { return h * 0x9E3779B97F4A7C15ull; }
static uint64_t rt_salt(uint64_t h) { return (h >> 16) & 0xffffu; }

static uint64_t rt_seek(const OracleJoin* j, uint64_t off, uint64_t salt, uint64_t* out_off)
{
    uint64_t e = j->entries[off];
    while (e != 0 && (e >> 48) != salt) {
        off = (off + 1) & j->mask;
        e = j->entries[off];
    }
    *out_off = off;
    return e;
}

/* rt_init + jr_build_rows: capacity pow2 >= max(1024, 2 rows); rows linked
 * from the last to the first so that chains list equal keys in build order. */
void oracle_join_build(OracleJoin* j)
{
    int64_t capacity = 1024;
    while (capacity < j->rows * 2) capacity *= 2;
    int shift = 64;
    for (int64_t c = capacity; c > 1; c >>= 1) --shift;
    j->entries = (uint64_t*)calloc((size_t)capacity, 8);
    j->capacity = capacity;
    j->mask = (uint64_t)(capacity - 1);
    j->shift = shift;
    j->keys = 0;
    for (int64_t k = j->rows - 1; k >= 0; --k) {
        uint64_t h = rt_mix(j->hash[k]);
        uint64_t salt = rt_salt(h), off;
        uint64_t e = rt_seek(j, h >> shift, salt, &off);
        for (;;) {
            if (e == 0) {
                j->next[k] = 0;
                j->keys++;
                j->entries[off] = (salt << 48) | (uint64_t)(k + 1);
                break;
            }
            int64_t head = (int64_t)(e & 0xffffffffffffull) - 1;
            if (key_equal(&j->key[head], &j->key[k], j->spec.NumKeys)) {
                j->next[k] = head + 1;
                j->entries[off] = (salt << 48) | (uint64_t)(k + 1);
                break;
            }
            off = (off + 1) & j->mask;
            e = rt_seek(j, off, salt, &off);
        }
    }
}

typedef struct PairBuf {
    int64_t* pairs;     /* 2 i64 each */
    int64_t n, cap;
} PairBuf;

/* jr_push_pair (qdb/kernel/join/join_rows.oz:117-125), pb_push x1.5 growth
 * (qdb/kernel/join/join_table.oz:51-74). */
static void push_pair(const OracleJoin* j, PairBuf* b, int64_t probe_row, int64_t build)
{
    if (b->n == b->cap) {
        b->cap = b->cap ? b->cap + b->cap / 2 : 4;
        b->pairs = (int64_t*)realloc(b->pairs, (size_t)b->cap * 16);
    }
    int64_t stream_id = (int64_t)((uint64_t)-1 << 32) + probe_row;
    if (j->spec.ProbeIsLeft) {
        b->pairs[2 * b->n] = stream_id;
        b->pairs[2 * b->n + 1] = build;
    } else {
        b->pairs[2 * b->n] = build;
        b->pairs[2 * b->n + 1] = stream_id;
    }
    b->n++;
}

/* Probe rows [r0, r1) of rs into b (the table is only read). */
static void probe_rows(const OracleJoin* j, const TRowSet* rs, int64_t r0, int64_t r1, PairBuf* b)
{
    const QdbRowJoinSpec* s = &j->spec;
    int outer = s->JoinType != QDB_JOIN_INNER;
    for (int64_t i = r0; i < r1; ++i) {
        if (rs->Selection && !rs->Selection[i]) continue;
        GroupKey key;
        int valid;
        join_read_key(&s->Probe, s->ProbeKeyColumns, s->NumKeys, rs, i, &key, &valid);
        if (!valid) {
            if (outer) push_pair(j, b, i, -1);
            continue;
        }
        int matched = 0;
        if (j->rows > 0) {
            uint64_t h = rt_mix(join_hash(&s->Probe, s->ProbeKeyColumns, s->NumKeys, rs, i));
            uint64_t salt = rt_salt(h), off;
            uint64_t e = rt_seek(j, h >> j->shift, salt, &off);
            while (e != 0) {
                int64_t head = (int64_t)(e & 0xffffffffffffull) - 1;
                if (key_equal(&j->key[head], &key, s->NumKeys)) {
                    for (int64_t cur = head + 1; cur != 0; cur = j->next[cur - 1]) {
                        push_pair(j, b, i, j->ordinal[cur - 1]);
                    }
                    matched = 1;
                    break;
                }
                off = (off + 1) & j->mask;
                e = rt_seek(j, off, salt, &off);
            }
        }
        if (!matched && outer) push_pair(j, b, i, -1);
    }
}

/* jr_probe_batch: pairs of (stream row id, build ordinal) in probe order;
 * a chain emits its rows in build order. Returns the pair count. The pair
 * buffer is reset by every probe (the processor drains it per batch). */
int64_t oracle_join_probe(OracleJoin* j, const TRowSet* rs)
{
    PairBuf b = {j->pairs, 0, j->pair_cap};
    probe_rows(j, rs, 0, rs->RowCount, &b);
    j->pairs = b.pairs;
    j->pair_cap = b.cap;
    j->npairs = b.n;
    return j->npairs;
}

/* The same probe over morsels of 262,144 rows (the reference's batch size,
 * qdb/io/parquet/source.cpp:333) taken by `threads` threads from a shared
 * counter, each into its own pair buffer; the buffers are joined in morsel
 * order, so the pairs equal oracle_join_probe's. For the CPU baseline. */
typedef struct ProbeWork {
    const OracleJoin* j;
    const TRowSet* rs;
    int64_t morsels, next;
    PairBuf* out;       /* one per morsel */
} ProbeWork;

enum { kMorselRows = 262144 };

static void* probe_worker(void* arg)
{
    ProbeWork* w = (ProbeWork*)arg;
    for (;;) {
        int64_t m = __atomic_fetch_add(&w->next, 1, __ATOMIC_RELAXED);
        if (m >= w->morsels) return NULL;
        int64_t r0 = m * kMorselRows, r1 = r0 + kMorselRows;
        if (r1 > w->rs->RowCount) r1 = w->rs->RowCount;
        probe_rows(w->j, w->rs, r0, r1, &w->out[m]);
    }
}

int64_t oracle_join_probe_parallel(OracleJoin* j, const TRowSet* rs, int threads)
{
    ProbeWork w;
    w.j = j;
    w.rs = rs;
    w.morsels = (rs->RowCount + kMorselRows - 1) / kMorselRows;
    w.next = 0;
    w.out = (PairBuf*)calloc((size_t)(w.morsels ? w.morsels : 1), sizeof(PairBuf));
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pthread_t tid[256];
    int started = 0;
    for (int t = 1; t < threads; ++t) {
        if (pthread_create(&tid[started], NULL, probe_worker, &w) == 0) started++;
    }
    probe_worker(&w);
    for (int t = 0; t < started; ++t) pthread_join(tid[t], NULL);
    int64_t total = 0;
    for (int64_t m = 0; m < w.morsels; ++m) total += w.out[m].n;
    if (total > j->pair_cap) {
        j->pair_cap = total;
        j->pairs = (int64_t*)realloc(j->pairs, (size_t)total * 16);
    }
    int64_t at = 0;
    for (int64_t m = 0; m < w.morsels; ++m) {
        if (w.out[m].n) memcpy(j->pairs + 2 * at, w.out[m].pairs, (size_t)w.out[m].n * 16);
        at += w.out[m].n;
        free(w.out[m].pairs);
    }
    free(w.out);
    j->npairs = total;
    return total;
}

const int64_t* oracle_join_pairs(const OracleJoin* j)
{
    return j->pairs;
}

int64_t oracle_join_keys(const OracleJoin* j)
{
    return j->keys;
}

int64_t oracle_join_rows(const OracleJoin* j)
{
    return j->rows;
}

/* ------------------------------------------------------------------------ */
/* Synthetic inputs on the host (the same generator the device runs).        */
/* ------------------------------------------------------------------------ */

#include "../include/qdb_gen.h"

void oracle_gen_lineitem(uint64_t seed, int64_t row0, int64_t rows, int8_t* flag, int8_t* status,
    int32_t* shipdate, int64_t* quantity, int64_t* price, int64_t* discount, int64_t* tax)
{
    for (int64_t i = 0; i < rows; ++i) {
        qdb_gen_lineitem_row(seed, row0 + i, &flag[i], &status[i], &shipdate[i], &quantity[i], &price[i], &discount[i], &tax[i]);
    }
}

void oracle_gen_join_build(uint64_t seed, int64_t row0, int64_t rows, int64_t build_rows, int64_t* key, int32_t* segment)
{
    for (int64_t i = 0; i < rows; ++i) qdb_gen_join_build_row(seed, row0 + i, build_rows, &key[i], &segment[i]);
}

void oracle_gen_join_probe(uint64_t seed, int64_t row0, int64_t rows, int64_t build_rows, int64_t* fk, int64_t* value)
{
    for (int64_t i = 0; i < rows; ++i) qdb_gen_join_probe_row(seed, row0 + i, build_rows, &fk[i], &value[i]);
}

void oracle_zipf_cdf(double* cdf, int64_t n)
{
    qdb_gen_zipf1_cdf(cdf, n);
}

void oracle_gen_groups(uint64_t seed, int64_t row0, int64_t rows, const double* cdf, int64_t distinct,
    int64_t* key, int64_t* vi, double* vf)
{
    for (int64_t i = 0; i < rows; ++i) qdb_gen_groups_row(seed, row0 + i, cdf, distinct, &key[i], &vi[i], &vf[i]);
}

/* ------------------------------------------------------------------------ */
/* Runtime join filter (qdb/kernel/join/runtime_filter.oz; policy in         */
/* qdb/exec/runtime_filter.cpp:122-300), for one producer                    */
/* ------------------------------------------------------------------------ */

/* rf_mix: the splitmix64 finalizer (runtime_filter.oz:79-86). */
uint64_t oracle_rf_mix(uint64_t v)
{
    v ^= v >> 30;
    v *= 0xBF58476D1CE4E5B9ull;
    v ^= v >> 27;
    v *= 0x94D049BB133111EBull;
    return v ^ (v >> 31);
}

typedef struct {
    int64_t max_keys;
    int64_t bloom_words;   /* bit_floor(bloom bytes) / 8, 0 below 64 bytes */
    uint64_t* keys;        /* every selected hash, in build order */
    int64_t nkeys, cap;
    int kind;              /* -1 building, 0 cannot answer, 1 exact, 2 Bloom */
    uint64_t* words;
    int64_t word_count;
    int64_t exact_keys;
} OracleRf;

void* oracle_rf_new(int64_t max_keys, int64_t bloom_bytes)
{
    OracleRf* f = (OracleRf*)calloc(1, sizeof(OracleRf));
    f->max_keys = max_keys;
    int64_t b = 0;
    if (bloom_bytes >= 64) { /* NormalizeBloomBytes, runtime_filter.cpp:20-22 */
        b = 1;
        while (b * 2 <= bloom_bytes) b *= 2;
    }
    f->bloom_words = b / 8;
    f->kind = -1;
    return f;
}

void oracle_rf_free(void* p)
{
    OracleRf* f = (OracleRf*)p;
    if (!f) return;
    free(f->keys);
    free(f->words);
    free(f);
}

/* rf_build: the selected hashes join the contribution. */
void oracle_rf_build(void* p, const uint64_t* hashes, const uint8_t* sel, int64_t count)
{
    OracleRf* f = (OracleRf*)p;
    for (int64_t i = 0; i < count; ++i) {
        if (sel && !sel[i]) continue;
        if (f->nkeys == f->cap) {
            f->cap = f->cap ? 2 * f->cap : 1024;
            f->keys = (uint64_t*)realloc(f->keys, (size_t)f->cap * 8);
        }
        f->keys[f->nkeys++] = hashes[i];
    }
}

static uint64_t rf_exact_stored(uint64_t h) { return h == 0 ? 1 : h; }

/* rf_exact_slot: the slot holding hash, or the empty slot where it belongs. */
static int64_t rf_exact_slot(const uint64_t* table, uint64_t mask, uint64_t h)
{
    const uint64_t stored = rf_exact_stored(h);
    int64_t slot = (int64_t)(h & mask);
    while (table[slot] != 0 && table[slot] != stored) slot = (int64_t)((uint64_t)(slot + 1) & mask);
    return slot;
}

static void rf_bloom_add(uint64_t* words, uint64_t block_mask, uint64_t h)
{
    const int64_t base = (int64_t)((oracle_rf_mix(h) & block_mask) * 8);
    const uint64_t bits = oracle_rf_mix(h ^ 0x9E3779B97F4A7C15ull);
    for (int w = 0; w < 8; ++w) words[base + w] |= 1ull << ((bits >> (8 * w)) & 63);
}

static int rf_bloom_contains(const uint64_t* words, uint64_t block_mask, uint64_t h)
{
    const int64_t base = (int64_t)((oracle_rf_mix(h) & block_mask) * 8);
    const uint64_t bits = oracle_rf_mix(h ^ 0x9E3779B97F4A7C15ull);
    for (int w = 0; w < 8; ++w) {
        if ((words[base + w] & (1ull << ((bits >> (8 * w)) & 63))) == 0) return 0;
    }
    return 1;
}

/* rf_publish (runtime_filter.oz:503-557): distinct hashes counted up to the
 * first one over MaxKeys; an exact table of pow2 >= 2 * distinct slots (at
 * least 8) filled in first-seen order, else every hash into the Bloom words. */
int oracle_rf_publish(void* p)
{
    OracleRf* f = (OracleRf*)p;
    if (f->kind >= 0) return f->kind;
    /* The set of distinct RAW hashes (the reference's deduplicating set
     * holds raw hashes, so 0 and 1 are two keys), first-seen order: a plain
     * open-addressing set with an occupancy byte per slot. */
    int64_t cap = 16;
    while (cap < 2 * (f->nkeys + 1)) cap *= 2;
    uint64_t* set = (uint64_t*)calloc((size_t)cap, 8);
    uint8_t* used = (uint8_t*)calloc((size_t)cap, 1);
    uint64_t* distinct = (uint64_t*)malloc((size_t)(f->nkeys ? f->nkeys : 1) * 8);
    int64_t nd = 0;
    for (int64_t i = 0; i < f->nkeys && nd <= f->max_keys; ++i) {
        const uint64_t h = f->keys[i];
        int64_t s = (int64_t)(oracle_rf_mix(h) & (uint64_t)(cap - 1));
        while (used[s] && set[s] != h) s = (s + 1) & (cap - 1);
        if (!used[s]) {
            used[s] = 1;
            set[s] = h;
            distinct[nd++] = h;
        }
    }
    free(set);
    free(used);
    if (nd <= f->max_keys) {
        int64_t size = 8;
        while (size < nd * 2) size *= 2;
        f->words = (uint64_t*)calloc((size_t)size, 8);
        for (int64_t i = 0; i < nd; ++i) {
            const int64_t s = rf_exact_slot(f->words, (uint64_t)(size - 1), distinct[i]);
            f->words[s] = rf_exact_stored(distinct[i]);
        }
        f->word_count = size;
        f->exact_keys = nd;
        f->kind = 1;
    } else if (f->bloom_words == 0) {
        f->kind = 0;
    } else {
        f->words = (uint64_t*)calloc((size_t)f->bloom_words, 8);
        f->word_count = f->bloom_words;
        for (int64_t i = 0; i < f->nkeys; ++i) rf_bloom_add(f->words, (uint64_t)(f->bloom_words / 8 - 1), f->keys[i]);
        f->kind = 2;
    }
    free(distinct);
    return f->kind;
}

/* rf_probe: output[row] = selected and (maybe) in the filter; kept rows. */
int64_t oracle_rf_probe(void* p, const uint64_t* hashes, const uint8_t* sel, uint8_t* out, int64_t count)
{
    const OracleRf* f = (const OracleRf*)p;
    int64_t kept = 0;
    for (int64_t i = 0; i < count; ++i) {
        int keep = !sel || sel[i];
        if (keep && f->kind == 1) {
            keep = f->words[rf_exact_slot(f->words, (uint64_t)(f->word_count - 1), hashes[i])] == rf_exact_stored(hashes[i]);
        } else if (keep && f->kind == 2) {
            keep = rf_bloom_contains(f->words, (uint64_t)(f->word_count / 8 - 1), hashes[i]);
        }
        out[i] = (uint8_t)keep;
        kept += keep;
    }
    return kept;
}
