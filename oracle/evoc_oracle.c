/*
 * ORACLE — test infrastructure only.
 *
 * A plain-C restatement of the reference's CPU kNN-graph path (eugenehp/evoc-rs
 * 0.0.1), used by the tests as the checker for the HIP library and by bench.py
 * as the timed CPU baseline.  The product library never links, loads or calls
 * anything in this directory.
 *
 * Every function cites the reference file:line it follows.  Where the
 * reference runs `n_threads` rayon workers, this code runs the same per-worker
 * loops (same block partition, same RNG streams, same bucket layout) and, when
 * more than one OS thread is allowed, executes the independent workers with
 * OpenMP.  Workers touch disjoint heap rows in every phase, so the result does
 * not depend on how the workers are scheduled, exactly as in the reference
 * (except for the reference's own update-buffer overflow spill, see
 * src/nndescent/float.rs:581-583, which is racy there and sequential here).
 *
 * Parity pins (tests/test_oracle.py):
 *   - NumPy RandomState KAT, src/numpy_rng.rs:232-252
 *   - fast_cosine KATs, src/fast_cosine.rs:70-85
 *   - RNG state after knn_graph, tests/parity_intermediates.rs:656-687
 *   - RP forest leaf array golden (medium_800/leaf_array.npy)
 *   - nn_inds / nn_dists goldens under tests/parity.rs:34-68
 *   - exact search vs brute force / numpy on small inputs
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_F32 0
#define ORC_I8 1
#define ORC_U8 2

#define ORC_INF FLT_MAX /* src/heap.rs:7 */
#define ORC_EPS 1e-8f   /* src/nndescent/float.rs:12 */

float orc_fast_cosine_c(const float *x, const float *y, int dim); /* oracle_fastmath.c */

/* Optional replacement for the deterministic f32 distance (the reference's own
 * native/fast_cosine.c compiled into oracle/_ref, when the tests load it). */
typedef float (*orc_cos_fn)(const float *, const float *, int);
static orc_cos_fn g_strict_cos = 0;
void orc_set_strict_cosine(orc_cos_fn fn) { g_strict_cos = fn; }

/* ------------------------------------------------------------------------- */
/* NumPy RandomState (MT19937 + legacy distributions): src/numpy_rng.rs:1-226 */
/* ------------------------------------------------------------------------- */
typedef struct {
    uint32_t key[624];
    int32_t pos;
    int32_t has_gauss;
    double gauss;
} orc_nprng;

/* src/numpy_rng.rs:35-45 */
void orc_nprng_seed(orc_nprng *r, uint32_t seed)
{
    r->key[0] = seed;
    for (uint32_t i = 1; i < 624; i++) {
        seed = 1812433253u * (seed ^ (seed >> 30)) + i;
        r->key[i] = seed;
    }
    r->pos = 624;
    r->has_gauss = 0;
    r->gauss = 0.0;
}

/* src/numpy_rng.rs:154-166 */
void orc_nprng_set_state(orc_nprng *r, const uint32_t *key, int32_t pos, int32_t has_gauss,
                         double gauss)
{
    memcpy(r->key, key, sizeof(r->key));
    r->pos = pos;
    r->has_gauss = has_gauss;
    r->gauss = gauss;
}

void orc_nprng_get_state(const orc_nprng *r, uint32_t *key, int32_t *pos)
{
    memcpy(key, r->key, sizeof(r->key));
    *pos = r->pos;
}

/* src/numpy_rng.rs:47-59 */
static void mt_gen(orc_nprng *r)
{
    uint32_t *k = r->key;
    int i;
    uint32_t y;
    for (i = 0; i < 624 - 397; i++) {
        y = (k[i] & 0x80000000u) | (k[i + 1] & 0x7fffffffu);
        k[i] = k[i + 397] ^ (y >> 1) ^ ((0u - (y & 1u)) & 0x9908b0dfu);
    }
    for (; i < 623; i++) {
        y = (k[i] & 0x80000000u) | (k[i + 1] & 0x7fffffffu);
        k[i] = k[i + 397 - 624] ^ (y >> 1) ^ ((0u - (y & 1u)) & 0x9908b0dfu);
    }
    y = (k[623] & 0x80000000u) | (k[0] & 0x7fffffffu);
    k[623] = k[396] ^ (y >> 1) ^ ((0u - (y & 1u)) & 0x9908b0dfu);
    r->pos = 0;
}

/* src/numpy_rng.rs:61-72 */
uint32_t orc_nprng_next_u32(orc_nprng *r)
{
    if (r->pos >= 624)
        mt_gen(r);
    uint32_t y = r->key[r->pos++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

/* src/numpy_rng.rs:74-78 */
double orc_nprng_random(orc_nprng *r)
{
    int32_t a = (int32_t)(orc_nprng_next_u32(r) >> 5);
    int32_t b = (int32_t)(orc_nprng_next_u32(r) >> 6);
    return ((double)a * 67108864.0 + (double)b) / 9007199254740992.0;
}

/* src/numpy_rng.rs:95-112 */
static uint64_t random_interval(orc_nprng *r, uint64_t max)
{
    if (max == 0)
        return 0;
    uint64_t mask = max;
    mask |= mask >> 1;
    mask |= mask >> 2;
    mask |= mask >> 4;
    mask |= mask >> 8;
    mask |= mask >> 16;
    mask |= mask >> 32;
    for (;;) {
        uint64_t v = (uint64_t)orc_nprng_next_u32(r) & mask;
        if (v <= max)
            return v;
    }
}

/* src/numpy_rng.rs:81-88 */
int64_t orc_nprng_randint(orc_nprng *r, int64_t low, int64_t high)
{
    uint64_t range = (uint64_t)(high - low);
    if (range == 0)
        return low;
    return low + (int64_t)random_interval(r, range - 1);
}

/* src/numpy_rng.rs:115-133 */
double orc_nprng_normal(orc_nprng *r)
{
    if (r->has_gauss) {
        r->has_gauss = 0;
        double t = r->gauss;
        r->gauss = 0.0;
        return t;
    }
    for (;;) {
        double x1 = 2.0 * orc_nprng_random(r) - 1.0;
        double x2 = 2.0 * orc_nprng_random(r) - 1.0;
        double r2 = x1 * x1 + x2 * x2;
        if (r2 < 1.0 && r2 != 0.0) {
            double f = sqrt(-2.0 * log(r2) / r2);
            r->gauss = f * x1;
            r->has_gauss = 1;
            return f * x2;
        }
    }
}

#define ORC_RI_LO ((int64_t)INT32_MIN + 1) /* src/knn.rs:11 */
#define ORC_RI_HI ((int64_t)INT32_MAX - 1) /* src/knn.rs:12 */

/* src/numpy_rng.rs:204-212 */
void orc_nprng_randint3_for_tau(orc_nprng *r, int64_t out[3])
{
    for (int i = 0; i < 3; i++)
        out[i] = orc_nprng_randint(r, ORC_RI_LO, ORC_RI_HI);
}

/* ------------------------------------------------------------------------- */
/* Tausworthe RNG: src/rng.rs:5-35                                            */
/* ------------------------------------------------------------------------- */
static inline int64_t shl64(int64_t v, int s) { return (int64_t)((uint64_t)v << s); }

/* src/rng.rs:5-13 */
int32_t orc_tau_rand_int(int64_t *s)
{
    s[0] = (shl64(s[0] & 4294967294LL, 12) & 0xFFFFFFFFLL) ^
           (((shl64(s[0], 13) & 0xFFFFFFFFLL) ^ s[0]) >> 19);
    s[1] = (shl64(s[1] & 4294967288LL, 4) & 0xFFFFFFFFLL) ^
           (((shl64(s[1], 2) & 0xFFFFFFFFLL) ^ s[1]) >> 25);
    s[2] = (shl64(s[2] & 4294967280LL, 17) & 0xFFFFFFFFLL) ^
           (((shl64(s[2], 3) & 0xFFFFFFFFLL) ^ s[2]) >> 11);
    return (int32_t)(uint32_t)(uint64_t)(s[0] ^ s[1] ^ s[2]);
}

/* i32::abs in a release build wraps at i32::MIN */
static inline int32_t wrap_abs(int32_t v) { return v == INT32_MIN ? v : (v < 0 ? -v : v); }

/* src/rng.rs:17-20 */
float orc_tau_rand(int64_t *s)
{
    int32_t v = orc_tau_rand_int(s);
    return (float)wrap_abs(v) / (float)0x7FFFFFFF;
}

/* src/rng.rs:24-26 (rem_euclid) */
static inline int32_t tau_rand_mod(int64_t *s, int32_t m)
{
    int32_t v = orc_tau_rand_int(s) % m;
    return v < 0 ? v + m : v;
}

/* src/rng.rs:29-35 */
static inline void offset_state(const int64_t *base, int64_t off, int64_t *out)
{
    for (int i = 0; i < 3; i++)
        out[i] = (int64_t)((uint64_t)base[i] + (uint64_t)off);
}

/* ------------------------------------------------------------------------- */
/* Distances                                                                  */
/* ------------------------------------------------------------------------- */
typedef struct {
    int dtype;
    const void *x;
    int64_t n, d; /* d = elements per row (bytes for u8) */
    int strict;   /* f32 only: strict => C fast-math dot (src/fast_cosine.rs:33-45) */
} orc_data;

/* x86_64 branch of dot_f32_rlx: serial, un-reassociated sum (src/fast_cosine.rs:55-62) */
float orc_dot_serial(const float *x, const float *y, int n)
{
    float s = 0.0f;
    for (int i = 0; i < n; i++)
        s += x[i] * y[i];
    return s;
}

/* src/fast_cosine.rs:33-45 */
static inline float cosine_f32(const orc_data *D, int64_t a, int64_t b)
{
    const float *X = (const float *)D->x;
    const float *xa = X + a * D->d, *xb = X + b * D->d;
    if (D->strict)
        return g_strict_cos ? g_strict_cos(xa, xb, (int)D->d) : orc_fast_cosine_c(xa, xb, (int)D->d);
    float dot = orc_dot_serial(xa, xb, (int)D->d);
    return dot > 0.0f ? -dot : FLT_MIN;
}

/* src/nndescent/int8.rs:15-21 */
static inline float dissim_i8(const orc_data *D, int64_t a, int64_t b)
{
    const int8_t *X = (const int8_t *)D->x;
    const int8_t *xa = X + a * D->d, *xb = X + b * D->d;
    int32_t acc = 0;
    for (int64_t i = 0; i < D->d; i++)
        acc += (int32_t)xa[i] * (int32_t)xb[i];
    return -(float)acc;
}

/* src/nndescent/uint8.rs:21-36 */
static inline float jaccard_u8(const orc_data *D, int64_t a, int64_t b)
{
    const uint8_t *X = (const uint8_t *)D->x;
    const uint8_t *xa = X + a * D->d, *xb = X + b * D->d;
    uint32_t inter = 0, uni = 0;
    for (int64_t i = 0; i < D->d; i++) {
        inter += (uint32_t)__builtin_popcount((unsigned)(xa[i] & xb[i]));
        uni += (uint32_t)__builtin_popcount((unsigned)(xa[i] | xb[i]));
    }
    return uni > 0 ? -((float)inter / (float)uni) : 0.0f;
}

static inline float dist(const orc_data *D, int64_t a, int64_t b)
{
    switch (D->dtype) {
    case ORC_F32: return cosine_f32(D, a, b);
    case ORC_I8: return dissim_i8(D, a, b);
    default: return jaccard_u8(D, a, b);
    }
}

float orc_distance(int dtype, const void *x, int64_t d, int strict, int64_t a, int64_t b)
{
    orc_data D = {dtype, x, 0, d, strict};
    return dist(&D, a, b);
}

/* ------------------------------------------------------------------------- */
/* Heaps: src/heap.rs                                                         */
/* ------------------------------------------------------------------------- */
typedef struct {
    int64_t n, k;
    int32_t *idx;
    float *dist;
    uint8_t *flag;
} orc_heap;

/* src/heap.rs:29-34 */
static int heap_make(orc_heap *h, int64_t n, int64_t k)
{
    h->n = n;
    h->k = k;
    h->idx = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n * k));
    h->dist = (float *)malloc(sizeof(float) * (size_t)(n * k));
    h->flag = (uint8_t *)calloc((size_t)(n * k), 1);
    if (!h->idx || !h->dist || !h->flag)
        return -1;
    for (int64_t i = 0; i < n * k; i++) {
        h->idx[i] = -1;
        h->dist[i] = ORC_INF;
    }
    return 0;
}

static void heap_free(orc_heap *h)
{
    free(h->idx);
    free(h->dist);
    free(h->flag);
}

/* src/heap.rs:143-200 */
static inline int flagged_heap_push(float *pri, int32_t *ind, uint8_t *flg, int64_t size, float p,
                                    int32_t nn)
{
    if (p >= pri[0])
        return 0;
    for (int64_t i = 0; i < size; i++)
        if (nn == ind[i])
            return 0;
    pri[0] = p;
    ind[0] = nn;
    int64_t i = 0;
    for (;;) {
        int64_t c1 = 2 * i + 1, c2 = c1 + 1, sw;
        if (c1 >= size)
            break;
        else if (c2 >= size) {
            if (pri[c1] > p)
                sw = c1;
            else
                break;
        } else if (pri[c1] >= pri[c2]) {
            if (p < pri[c1])
                sw = c1;
            else
                break;
        } else if (p < pri[c2])
            sw = c2;
        else
            break;
        pri[i] = pri[sw];
        ind[i] = ind[sw];
        flg[i] = flg[sw];
        i = sw;
    }
    pri[i] = p;
    ind[i] = nn;
    flg[i] = 1;
    return 1;
}

/* src/heap.rs:86-140 */
static inline int cand_heap_push(float *pri, int32_t *ind, int64_t size, float p, int32_t nn)
{
    if (p >= pri[0])
        return 0;
    for (int64_t i = 0; i < size; i++)
        if (nn == ind[i])
            return 0;
    pri[0] = p;
    ind[0] = nn;
    int64_t i = 0;
    for (;;) {
        int64_t c1 = 2 * i + 1, c2 = c1 + 1, sw;
        if (c1 >= size)
            break;
        else if (c2 >= size) {
            if (pri[c1] > p)
                sw = c1;
            else
                break;
        } else if (pri[c1] >= pri[c2]) {
            if (p < pri[c1])
                sw = c1;
            else
                break;
        } else if (p < pri[c2])
            sw = c2;
        else
            break;
        pri[i] = pri[sw];
        ind[i] = ind[sw];
        i = sw;
    }
    pri[i] = p;
    ind[i] = nn;
    return 1;
}

/* src/heap.rs:37-60 */
static void siftdown(float *h1, int32_t *h2, int64_t n, int64_t elt)
{
    while (elt * 2 + 1 < n) {
        int64_t l = elt * 2 + 1, r = l + 1, sw = elt;
        if (h1[sw] < h1[l])
            sw = l;
        if (r < n && h1[sw] < h1[r])
            sw = r;
        if (sw == elt)
            break;
        float tf = h1[elt];
        h1[elt] = h1[sw];
        h1[sw] = tf;
        int32_t ti = h2[elt];
        h2[elt] = h2[sw];
        h2[sw] = ti;
        elt = sw;
    }
}

/* src/heap.rs:62-83 */
void orc_deheap_sort(int32_t *idx, float *dst, int64_t n, int64_t k, int nw)
{
    (void)nw;
#pragma omp parallel for num_threads(nw) schedule(static) if (nw > 1)
    for (int64_t r = 0; r < n; r++) {
        int32_t *ir = idx + r * k;
        float *dr = dst + r * k;
        for (int64_t j = k - 1; j >= 1; j--) {
            int32_t ti = ir[0];
            ir[0] = ir[j];
            ir[j] = ti;
            float tf = dr[0];
            dr[0] = dr[j];
            dr[j] = tf;
            siftdown(dr, ir, j, 0);
        }
    }
}

/* src/heap.rs:203-343.  Workers n in [0, T) own vertex block [n*bs, (n+1)*bs). */
static int build_candidates(orc_heap *g, int64_t mc, const int64_t *rng_state, int64_t T, int nw,
                            int32_t *new_idx, int32_t *old_idx)
{
    const int64_t N = g->n, K = g->k;
    float *new_pri = (float *)malloc(sizeof(float) * (size_t)(N * mc));
    float *old_pri = (float *)malloc(sizeof(float) * (size_t)(N * mc));
    if (!new_pri || !old_pri) {
        free(new_pri);
        free(old_pri);
        return -1;
    }
    for (int64_t i = 0; i < N * mc; i++) {
        new_idx[i] = -1;
        old_idx[i] = -1;
        new_pri[i] = ORC_INF;
        old_pri[i] = ORC_INF;
    }
    const int64_t bs = N / T + 1;
    (void)nw;
#pragma omp parallel for num_threads(nw) schedule(dynamic, 1) if (nw > 1)
    for (int64_t n = 0; n < T; n++) {
        int64_t st[3];
        offset_state(rng_state, n, st);
        const int64_t b0 = n * bs, b1 = (b0 + bs < N) ? b0 + bs : N;
        for (int64_t i = 0; i < N; i++) {
            for (int64_t j = 0; j < K; j++) {
                int32_t idx = g->idx[i * K + j];
                if (idx < 0)
                    continue;
                int in_i = i >= b0 && i < b1;
                int in_j = (int64_t)idx >= b0 && (int64_t)idx < b1;
                if (!in_i && !in_j)
                    continue;
                int isn = g->flag[i * K + j] != 0;
                float d = orc_tau_rand(st);
                float *pri = isn ? new_pri : old_pri;
                int32_t *ind = isn ? new_idx : old_idx;
                if (in_i)
                    cand_heap_push(pri + i * mc, ind + i * mc, mc, d, idx);
                if (in_j)
                    cand_heap_push(pri + (int64_t)idx * mc, ind + (int64_t)idx * mc, mc, d,
                                   (int32_t)i);
            }
        }
    }
    /* clear flags of graph entries that were sampled into the new lists: heap.rs:319-340 */
#pragma omp parallel for num_threads(nw) schedule(static) if (nw > 1)
    for (int64_t i = 0; i < N; i++) {
        for (int64_t j = 0; j < K; j++) {
            int32_t idx = g->idx[i * K + j];
            if (idx < 0)
                continue;
            for (int64_t c = 0; c < mc; c++)
                if (new_idx[i * mc + c] == idx) {
                    g->flag[i * K + j] = 0;
                    break;
                }
        }
    }
    free(new_pri);
    free(old_pri);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Random-projection forest: src/nndescent/{float,int8,uint8}.rs              */
/* ------------------------------------------------------------------------- */
typedef struct {
    int32_t *pts; /* concatenated leaf members */
    int64_t len, cap;
    int64_t *leaf_off; /* start offsets of leaves, n_leaves+1 */
    int64_t n_leaves, leaf_cap;
} orc_leaves;

static int leaves_push(orc_leaves *L, const int32_t *ix, int64_t m)
{
    if (L->len + m > L->cap) {
        int64_t nc = (L->cap ? L->cap * 2 : 1024);
        while (nc < L->len + m)
            nc *= 2;
        int32_t *np = (int32_t *)realloc(L->pts, sizeof(int32_t) * (size_t)nc);
        if (!np)
            return -1;
        L->pts = np;
        L->cap = nc;
    }
    if (L->n_leaves + 2 > L->leaf_cap) {
        int64_t nc = L->leaf_cap ? L->leaf_cap * 2 : 256;
        int64_t *no = (int64_t *)realloc(L->leaf_off, sizeof(int64_t) * (size_t)nc);
        if (!no)
            return -1;
        L->leaf_off = no;
        L->leaf_cap = nc;
        if (L->n_leaves == 0)
            L->leaf_off[0] = 0;
    }
    memcpy(L->pts + L->len, ix, sizeof(int32_t) * (size_t)m);
    L->len += m;
    L->n_leaves += 1;
    L->leaf_off[L->n_leaves] = L->len;
    return 0;
}

/* Hyperplane margin of row `pt` for the three split kinds. */
typedef struct {
    float *hv;           /* f32 / i8 hyperplane, d floats */
    uint8_t *pos, *neg;  /* u8 masks, d bytes */
} orc_plane;

/* float.rs:13-95 / int8.rs:24-123 / uint8.rs:38-111: returns n_left, fills tl / tr, and the
 * two pivot rows in piv[0] (left) and piv[1] (right) when piv is not NULL */
static int64_t rp_split(const orc_data *D, const int32_t *ind, int64_t m, int64_t *st,
                        orc_plane *pl, int32_t *tl, int32_t *tr, int64_t *n_right, int64_t *piv)
{
    const int64_t dim = D->d;
    uint32_t isz = (uint32_t)m;
    uint32_t li = (uint32_t)tau_rand_mod(st, (int32_t)isz);
    uint32_t ri = (uint32_t)tau_rand_mod(st, (int32_t)isz);
    ri += (li == ri);
    ri %= isz;
    int64_t left = ind[li], right = ind[ri];
    if (piv) {
        piv[0] = left;
        piv[1] = right;
    }
    if (D->dtype == ORC_F32) {
        const float *X = (const float *)D->x;
        const float *a = X + left * dim, *b = X + right * dim;
        float nrm = 0.0f;
        for (int64_t k = 0; k < dim; k++) {
            pl->hv[k] = a[k] - b[k];
            nrm += pl->hv[k] * pl->hv[k];
        }
        nrm = sqrtf(nrm);
        if (fabsf(nrm) < ORC_EPS)
            nrm = 1.0f;
        for (int64_t k = 0; k < dim; k++)
            pl->hv[k] /= nrm;
    } else if (D->dtype == ORC_I8) {
        const int8_t *X = (const int8_t *)D->x;
        const int8_t *a = X + left * dim, *b = X + right * dim;
        float ln = 0.0f, rn = 0.0f;
        for (int64_t k = 0; k < dim; k++) {
            float lv = (float)a[k], rv = (float)b[k];
            ln += lv * lv;
            rn += rv * rv;
        }
        ln = sqrtf(ln);
        rn = sqrtf(rn);
        float nrm = 0.0f;
        for (int64_t k = 0; k < dim; k++) {
            float lv = ln > 0.0f ? (float)a[k] / ln : 0.0f;
            float rv = rn > 0.0f ? (float)b[k] / rn : 0.0f;
            pl->hv[k] = lv - rv;
            nrm += pl->hv[k] * pl->hv[k];
        }
        nrm = sqrtf(nrm);
        if (fabsf(nrm) < ORC_EPS)
            nrm = 1.0f;
        for (int64_t k = 0; k < dim; k++)
            pl->hv[k] /= nrm;
    } else {
        const uint8_t *X = (const uint8_t *)D->x;
        const uint8_t *a = X + left * dim, *b = X + right * dim;
        for (int64_t k = 0; k < dim; k++) {
            uint8_t xr = a[k] ^ b[k];
            pl->pos[k] = xr & a[k];
            pl->neg[k] = xr & b[k];
        }
    }
    int64_t nl = 0, nr = 0;
    for (int64_t t = 0; t < m; t++) {
        int64_t lst[3];
        offset_state(st, t, lst);
        int64_t p = ind[t];
        float margin = 0.0f;
        if (D->dtype == ORC_F32) {
            const float *x = (const float *)D->x + p * dim;
            for (int64_t k = 0; k < dim; k++)
                margin += pl->hv[k] * x[k];
        } else if (D->dtype == ORC_I8) {
            const int8_t *x = (const int8_t *)D->x + p * dim;
            for (int64_t k = 0; k < dim; k++)
                margin += pl->hv[k] * (float)x[k];
        } else {
            const uint8_t *x = (const uint8_t *)D->x + p * dim;
            for (int64_t k = 0; k < dim; k++) {
                margin += (float)__builtin_popcount((unsigned)(pl->pos[k] & x[k]));
                margin -= (float)__builtin_popcount((unsigned)(pl->neg[k] & x[k]));
            }
        }
        int cls;
        if (fabsf(margin) < ORC_EPS)
            cls = tau_rand_mod(lst, 2);
        else if (margin > 0.0f)
            cls = 0;
        else
            cls = 1;
        if (cls == 0)
            tl[nl++] = (int32_t)p;
        else
            tr[nr++] = (int32_t)p;
    }
    if (nl == 0 || nr == 0) {
        nl = nr = 0;
        for (int64_t t = 0; t < m; t++) {
            int cls = tau_rand_mod(st, 2);
            if (cls == 0)
                tl[nl++] = ind[t];
            else
                tr[nr++] = ind[t];
        }
    }
    *n_right = nr;
    return nl;
}

/* float.rs:97-126 (recursive depth-first, left subtree first) */
static int make_tree(const orc_data *D, const int32_t *ind, int64_t m, orc_leaves *L, int64_t *st,
                     int64_t leaf_size, int64_t max_depth, orc_plane *pl)
{
    if (m > leaf_size && max_depth > 0) {
        int32_t *tl = (int32_t *)malloc(sizeof(int32_t) * (size_t)m);
        int32_t *tr = (int32_t *)malloc(sizeof(int32_t) * (size_t)m);
        if (!tl || !tr) {
            free(tl);
            free(tr);
            return -1;
        }
        int64_t nr = 0;
        int64_t nl = rp_split(D, ind, m, st, pl, tl, tr, &nr, NULL);
        int rc = make_tree(D, tl, nl, L, st, leaf_size, max_depth - 1, pl);
        free(tl);
        if (rc == 0)
            rc = make_tree(D, tr, nr, L, st, leaf_size, max_depth - 1, pl);
        free(tr);
        return rc;
    }
    return leaves_push(L, ind, m);
}

/* float.rs:128-160: one tree -> leaf list (kept ragged; padded when concatenated) */
static int make_leaf_list(const orc_data *D, int64_t *st, int64_t leaf_size, int64_t max_depth,
                          orc_leaves *L)
{
    int32_t *ind = (int32_t *)malloc(sizeof(int32_t) * (size_t)D->n);
    orc_plane pl;
    pl.hv = (float *)malloc(sizeof(float) * (size_t)(D->d ? D->d : 1));
    pl.pos = (uint8_t *)malloc((size_t)(D->d ? D->d : 1));
    pl.neg = (uint8_t *)malloc((size_t)(D->d ? D->d : 1));
    int rc = -1;
    if (ind && pl.hv && pl.pos && pl.neg) {
        for (int64_t i = 0; i < D->n; i++)
            ind[i] = (int32_t)i;
        rc = make_tree(D, ind, D->n, L, st, leaf_size, max_depth, &pl);
    }
    free(ind);
    free(pl.hv);
    free(pl.pos);
    free(pl.neg);
    return rc;
}

/*
 * src/knn.rs:115-154 (make_forest) with the per-tree builders of
 * float.rs:162-177 / int8.rs:191-204 / uint8.rs:179-192.  Draws 3*n_trees
 * randint values from `rng` (src/knn.rs:124-129).  Returns a malloc'd
 * (rows x cols) leaf array padded with -1, or NULL with *rows = 0 when
 * n_trees == 0.  Each tree's leaf array is at least leaf_size wide
 * (float.rs:148-151), and the concatenation is as wide as the widest tree.
 */
int32_t *orc_make_forest(int dtype, const void *x, int64_t n, int64_t d, int64_t n_neighbors,
                         int64_t n_trees, int64_t leaf_size, orc_nprng *rng, int64_t max_depth,
                         int64_t *rows, int64_t *cols, int nw)
{
    leaf_size = leaf_size > 10 ? leaf_size : 10;
    leaf_size = leaf_size > n_neighbors ? leaf_size : n_neighbors;
    *rows = 0;
    *cols = 0;
    if (n_trees <= 0)
        return NULL;
    int64_t *states = (int64_t *)malloc(sizeof(int64_t) * 3 * (size_t)n_trees);
    orc_leaves *Ls = (orc_leaves *)calloc((size_t)n_trees, sizeof(orc_leaves));
    if (!states || !Ls) {
        free(states);
        free(Ls);
        return NULL;
    }
    for (int64_t t = 0; t < n_trees; t++)
        for (int j = 0; j < 3; j++)
            states[t * 3 + j] = orc_nprng_randint(rng, ORC_RI_LO, ORC_RI_HI);
    orc_data D = {dtype, x, n, d, 0};
    int fail = 0;
    (void)nw;
#pragma omp parallel for num_threads(nw) schedule(dynamic, 1) if (nw > 1) reduction(| : fail)
    for (int64_t t = 0; t < n_trees; t++)
        fail |= make_leaf_list(&D, states + t * 3, leaf_size, max_depth, &Ls[t]) != 0;
    int32_t *out = NULL;
    if (!fail) {
        int64_t total_rows = 0, maxw = 0;
        for (int64_t t = 0; t < n_trees; t++) {
            int64_t w = leaf_size;
            for (int64_t l = 0; l < Ls[t].n_leaves; l++) {
                int64_t s = Ls[t].leaf_off[l + 1] - Ls[t].leaf_off[l];
                if (s > w)
                    w = s;
            }
            if (w > maxw)
                maxw = w;
            total_rows += Ls[t].n_leaves;
        }
        out = (int32_t *)malloc(sizeof(int32_t) * (size_t)(total_rows * maxw + 1));
        if (out) {
            for (int64_t i = 0; i < total_rows * maxw; i++)
                out[i] = -1;
            int64_t r = 0;
            for (int64_t t = 0; t < n_trees; t++)
                for (int64_t l = 0; l < Ls[t].n_leaves; l++, r++) {
                    int64_t s = Ls[t].leaf_off[l + 1] - Ls[t].leaf_off[l];
                    memcpy(out + r * maxw, Ls[t].pts + Ls[t].leaf_off[l], sizeof(int32_t) * (size_t)s);
                }
            *rows = total_rows;
            *cols = maxw;
        }
    }
    for (int64_t t = 0; t < n_trees; t++) {
        free(Ls[t].pts);
        free(Ls[t].leaf_off);
    }
    free(Ls);
    free(states);
    return out;
}

void orc_free(void *p) { free(p); }

/*
 * One random-projection split as make_tree takes it (float.rs:13-95 / int8.rs:24-123 /
 * uint8.rs:38-111), for the tests: rows ind[0..m) of the n x d matrix, the tau state st
 * (advanced in place), left members then right members into out (m entries), the two
 * pivot rows into piv[2].  Returns the left count, or -1 if scratch cannot be allocated.
 */
int64_t orc_rp_split(int dtype, const void *x, int64_t n, int64_t d, const int32_t *ind, int64_t m,
                     int64_t *st, int32_t *out, int64_t *piv)
{
    orc_data D = {dtype, x, n, d, 0};
    orc_plane pl;
    pl.hv = (float *)malloc(sizeof(float) * (size_t)(d > 0 ? d : 1));
    pl.pos = (uint8_t *)malloc((size_t)(d > 0 ? d : 1));
    pl.neg = (uint8_t *)malloc((size_t)(d > 0 ? d : 1));
    int32_t *tr = (int32_t *)malloc(sizeof(int32_t) * (size_t)(m > 0 ? m : 1));
    int64_t nl = -1;
    if (pl.hv && pl.pos && pl.neg && tr && m >= 2) {
        int64_t nr = 0;
        nl = rp_split(&D, ind, m, st, &pl, out, tr, &nr, piv);
        memcpy(out + nl, tr, sizeof(int32_t) * (size_t)nr);
    }
    free(pl.hv);
    free(pl.pos);
    free(pl.neg);
    free(tr);
    return nl;
}

/* ------------------------------------------------------------------------- */
/* NN-descent initialisation                                                  */
/* ------------------------------------------------------------------------- */
typedef struct {
    int32_t p, q;
    float d;
} orc_upd;

/* float.rs:179-348 (init_rp_tree_float + generate_leaf_updates_float); same in int8/uint8 */
static int init_rp_tree(const orc_data *D, orc_heap *g, const int32_t *leaf, int64_t n_leaves,
                        int64_t max_leaf, int64_t T, int nw)
{
    const int64_t block = T * 64;
    const int64_t n_blocks = n_leaves / block;
    const int64_t N = g->n, K = g->k;
    const int64_t vbs = N / T + 1;
    /* per-worker capacity: every leaf row in its share emits self + all pairs */
    const int64_t rows_per_w = block / T + 1;
    const int64_t cap = rows_per_w * (max_leaf + max_leaf * (max_leaf - 1) / 2) + 1;
    orc_upd *buf = (orc_upd *)malloc(sizeof(orc_upd) * (size_t)(cap * T));
    int64_t *cnt = (int64_t *)calloc((size_t)T, sizeof(int64_t));
    float *thr = (float *)malloc(sizeof(float) * (size_t)N);
    if (!buf || !cnt || !thr) {
        free(buf);
        free(cnt);
        free(thr);
        return -1;
    }
    for (int64_t b = 0; b <= n_blocks; b++) {
        int64_t bs0 = b * block, bs1 = (b + 1) * block < n_leaves ? (b + 1) * block : n_leaves;
        if (bs0 >= bs1)
            continue;
        int64_t brows = bs1 - bs0;
        for (int64_t v = 0; v < N; v++)
            thr[v] = g->dist[v * K];
        int64_t rpt = brows / T + 1;
#pragma omp parallel for num_threads(nw) schedule(dynamic, 1) if (nw > 1)
        for (int64_t t = 0; t < T; t++) {
            int64_t c = 0;
            orc_upd *u = buf + t * cap;
            for (int64_t r = 0; r < rpt; r++) {
                int64_t row = t * rpt + r;
                if (row >= brows)
                    break;
                const int32_t *lr = leaf + (bs0 + row) * max_leaf;
                for (int64_t i = 0; i < max_leaf; i++) {
                    int32_t p = lr[i];
                    if (p < 0)
                        break;
                    u[c].p = p;
                    u[c].q = p;
                    u[c].d = -1.0f;
                    c++;
                    for (int64_t j = i + 1; j < max_leaf; j++) {
                        int32_t q = lr[j];
                        if (q < 0)
                            break;
                        float d = dist(D, p, q);
                        float mt = thr[p] > thr[q] ? thr[p] : thr[q];
                        if (d < mt) {
                            u[c].p = p;
                            u[c].q = q;
                            u[c].d = d;
                            c++;
                        }
                    }
                }
            }
            cnt[t] = c;
        }
#pragma omp parallel for num_threads(nw) schedule(dynamic, 1) if (nw > 1)
        for (int64_t t = 0; t < T; t++) {
            int64_t v0 = t * vbs, v1 = v0 + vbs < N ? v0 + vbs : N;
            for (int64_t j = 0; j < T; j++) {
                const orc_upd *u = buf + j * cap;
                for (int64_t e = 0; e < cnt[j]; e++) {
                    int32_t p = u[e].p, q = u[e].q;
                    if (p == -1)
                        continue;
                    float d = u[e].d;
                    if (p >= v0 && p < v1)
                        flagged_heap_push(g->dist + (int64_t)p * K, g->idx + (int64_t)p * K,
                                          g->flag + (int64_t)p * K, K, d, q);
                    if (q >= v0 && q < v1)
                        flagged_heap_push(g->dist + (int64_t)q * K, g->idx + (int64_t)q * K,
                                          g->flag + (int64_t)q * K, K, d, p);
                }
            }
        }
        memset(cnt, 0, sizeof(int64_t) * (size_t)T);
    }
    free(buf);
    free(cnt);
    free(thr);
    return 0;
}

/* float.rs:350-390 */
static void init_random(const orc_data *D, orc_heap *g, int64_t n_neighbors,
                        const int64_t *rng_state, int nw)
{
    const int64_t N = D->n, K = g->k;
    (void)nw;
#pragma omp parallel for num_threads(nw) schedule(static) if (nw > 1)
    for (int64_t i = 0; i < N; i++) {
        int64_t st[3];
        offset_state(rng_state, i, st);
        int32_t *ir = g->idx + i * K;
        if (ir[0] >= 0)
            continue;
        int64_t filled = 0;
        for (int64_t j = 0; j < K; j++)
            filled += ir[j] >= 0;
        int64_t to_fill = n_neighbors > filled ? n_neighbors - filled : 0;
        for (int64_t s = 0; s < to_fill; s++) {
            int32_t a = wrap_abs(orc_tau_rand_int(st));
            uint64_t idx = (uint64_t)(int64_t)a % (uint64_t)N;
            int dup = 0;
            for (int64_t j = 0; j < K; j++)
                if (ir[j] == (int32_t)idx) {
                    dup = 1;
                    break;
                }
            if (dup)
                continue;
            float d = dist(D, (int64_t)idx, i);
            flagged_heap_push(g->dist + i * K, ir, g->flag + i * K, K, d, (int32_t)idx);
        }
    }
}

/* ------------------------------------------------------------------------- */
/* NN-descent local join (sorted updates): float.rs:517-669, heap.rs:425-502  */
/* ------------------------------------------------------------------------- */
typedef struct {
    float *arr;     /* (T, max_updates, 3) */
    int64_t max_updates;
    int32_t *counts; /* (T, T+1) */
} orc_sorted_buf;

static inline void emit_sorted(orc_sorted_buf *B, int64_t T, int64_t t, int64_t blk,
                               int32_t *lc, int32_t p, int32_t q, float d)
{
    const int64_t mps = B->max_updates / T;
    int64_t w = t * mps + lc[blk];
    if (w < B->max_updates) {
        float *e = B->arr + (blk * B->max_updates + w) * 3;
        e[0] = (float)p;
        e[1] = (float)q;
        e[2] = d;
        lc[blk] += 1;
    }
}

static void gen_sorted_updates(const orc_data *D, orc_sorted_buf *B, const int32_t *newc,
                               const int32_t *oldc, int64_t brows, int64_t mcn, int64_t mco,
                               const float *thr, int64_t T, int nw)
{
    const int64_t rpt = brows / T + 1;
    const int64_t vbs = D->n / T + 1;
    memset(B->counts, 0, sizeof(int32_t) * (size_t)(T * (T + 1)));
#pragma omp parallel for num_threads(nw) schedule(dynamic, 1) if (nw > 1)
    for (int64_t t = 0; t < T; t++) {
        int32_t *lc = (int32_t *)calloc((size_t)T, sizeof(int32_t));
        for (int64_t r = 0; r < rpt; r++) {
            int64_t i = t * rpt + r;
            if (i >= brows)
                break;
            for (int64_t j = 0; j < mcn; j++) {
                int32_t p = newc[i * mcn + j];
                if (p < 0)
                    continue;
                float tp = thr[p];
                int64_t pb = p / vbs;
                if (pb >= T)
                    pb = T - 1;
                for (int64_t k = j + 1; k < mcn; k++) {
                    int32_t q = newc[i * mcn + k];
                    if (q < 0)
                        continue;
                    float d = dist(D, p, q);
                    float tq = thr[q];
                    float mt = tp > tq ? tp : tq;
                    if (d <= mt) {
                        int64_t qb = q / vbs;
                        if (qb >= T)
                            qb = T - 1;
                        emit_sorted(B, T, t, pb, lc, p, q, d);
                        if (qb != pb)
                            emit_sorted(B, T, t, qb, lc, p, q, d);
                    }
                }
                for (int64_t k = 0; k < mco; k++) {
                    int32_t q = oldc[i * mco + k];
                    if (q < 0)
                        continue;
                    float d = dist(D, p, q);
                    float tq = thr[q];
                    float mt = tp > tq ? tp : tq;
                    if (d <= mt) {
                        int64_t qb = q / vbs;
                        if (qb >= T)
                            qb = T - 1;
                        emit_sorted(B, T, t, pb, lc, p, q, d);
                        if (qb != pb)
                            emit_sorted(B, T, t, qb, lc, p, q, d);
                    }
                }
            }
        }
        for (int64_t b = 0; b < T; b++)
            B->counts[b * (T + 1) + t + 1] = lc[b];
        free(lc);
    }
}

/* heap.rs:425-502 */
static uint64_t apply_sorted_updates(orc_heap *g, const orc_sorted_buf *B, int64_t T, int nw)
{
    const int64_t N = g->n, K = g->k;
    const int64_t vbs = N / T + 1;
    const int64_t mps = B->max_updates / T;
    uint64_t total = 0;
#pragma omp parallel for num_threads(nw) schedule(dynamic, 1) if (nw > 1) reduction(+ : total)
    for (int64_t n = 0; n < T; n++) {
        int64_t v0 = n * vbs, v1 = v0 + vbs < N ? v0 + vbs : N;
        uint32_t local = 0;
        for (int64_t t = 0; t < T; t++) {
            int64_t start = t * mps;
            int64_t count = B->counts[n * (T + 1) + t + 1];
            for (int64_t j = 0; j < count; j++) {
                const float *e = B->arr + (n * B->max_updates + start + j) * 3;
                int32_t p = (int32_t)e[0], q = (int32_t)e[1];
                float d = e[2];
                if (p >= v0 && p < v1)
                    local += (uint32_t)flagged_heap_push(g->dist + (int64_t)p * K,
                                                         g->idx + (int64_t)p * K,
                                                         g->flag + (int64_t)p * K, K, d, q);
                if (q >= v0 && q < v1)
                    local += (uint32_t)flagged_heap_push(g->dist + (int64_t)q * K,
                                                         g->idx + (int64_t)q * K,
                                                         g->flag + (int64_t)q * K, K, d, p);
            }
        }
        total += local;
    }
    return total;
}

/* ------------------------------------------------------------------------- */
/* Unsorted local join (f32 only): float.rs:393-514, heap.rs:346-422          */
/* ------------------------------------------------------------------------- */
static void gen_unsorted_updates(const orc_data *D, float *arr, int64_t max_updates,
                                 int32_t *cnt, const int32_t *newc, const int32_t *oldc,
                                 int64_t brows, int64_t mcn, int64_t mco, const float *thr,
                                 int64_t T, int nw)
{
    const int64_t rpt = brows / T + 1;
    const int64_t ws = 8;
#pragma omp parallel for num_threads(nw) schedule(dynamic, 1) if (nw > 1)
    for (int64_t t = 0; t < T; t++) {
        int64_t idx = 0;
        float *u = arr + t * max_updates * 3;
#define ORC_EMIT(P, Q, DD)                                                                        \
    do {                                                                                          \
        u[idx * 3] = (float)(P);                                                                  \
        u[idx * 3 + 1] = (float)(Q);                                                              \
        u[idx * 3 + 2] = (DD);                                                                    \
        idx++;                                                                                    \
    } while (0)
        for (int64_t r = 0; r < rpt; r++) {
            int64_t i = t * rpt + r;
            if (i >= brows || idx >= max_updates)
                break;
            int64_t ns = 0;
            while (ns < mcn && idx < max_updates) {
                int64_t ne = ns + ws < mcn ? ns + ws : mcn;
                for (int64_t j = ns; j < ne; j++) {
                    if (idx >= max_updates)
                        break;
                    int32_t p = newc[i * mcn + j];
                    if (p < 0)
                        continue;
                    float tp = thr[p];
                    for (int64_t k = j + 1; k < ne; k++) {
                        if (idx >= max_updates)
                            break;
                        int32_t q = newc[i * mcn + k];
                        if (q < 0)
                            continue;
                        float d = dist(D, p, q);
                        float mt = tp > thr[q] ? tp : thr[q];
                        if (d <= mt)
                            ORC_EMIT(p, q, d);
                    }
                    for (int64_t k = ne; k < mcn; k++) {
                        if (idx >= max_updates)
                            break;
                        int32_t q = newc[i * mcn + k];
                        if (q < 0)
                            continue;
                        float d = dist(D, p, q);
                        float mt = tp > thr[q] ? tp : thr[q];
                        if (d <= mt)
                            ORC_EMIT(p, q, d);
                    }
                    int64_t os = 0;
                    while (os < mco && idx < max_updates) {
                        int64_t oe = os + ws < mco ? os + ws : mco;
                        for (int64_t k = os; k < oe; k++) {
                            if (idx >= max_updates)
                                break;
                            int32_t q = oldc[i * mco + k];
                            if (q < 0)
                                continue;
                            float d = dist(D, p, q);
                            float mt = tp > thr[q] ? tp : thr[q];
                            if (d <= mt)
                                ORC_EMIT(p, q, d);
                        }
                        os = oe;
                    }
                }
                ns = ne;
            }
        }
#undef ORC_EMIT
        cnt[t] = (int32_t)idx;
    }
}

static uint64_t apply_unsorted_updates(orc_heap *g, const float *arr, int64_t max_updates,
                                       const int32_t *cnt, int64_t T, int nw)
{
    const int64_t N = g->n, K = g->k;
    const int64_t vbs = N / T + 1;
    uint64_t total = 0;
#pragma omp parallel for num_threads(nw) schedule(dynamic, 1) if (nw > 1) reduction(+ : total)
    for (int64_t n = 0; n < T; n++) {
        int64_t v0 = n * vbs, v1 = v0 + vbs < N ? v0 + vbs : N;
        uint32_t local = 0;
        for (int64_t t = 0; t < T; t++) {
            for (int64_t j = 0; j < cnt[t]; j++) {
                const float *e = arr + (t * max_updates + j) * 3;
                int32_t p = (int32_t)e[0];
                if (p == -1)
                    break;
                int32_t q = (int32_t)e[1];
                float d = e[2];
                if (p >= v0 && p < v1)
                    local += (uint32_t)flagged_heap_push(g->dist + (int64_t)p * K,
                                                         g->idx + (int64_t)p * K,
                                                         g->flag + (int64_t)p * K, K, d, q);
                if (q >= v0 && q < v1)
                    local += (uint32_t)flagged_heap_push(g->dist + (int64_t)q * K,
                                                         g->idx + (int64_t)q * K,
                                                         g->flag + (int64_t)q * K, K, d, p);
            }
        }
        total += local;
    }
    return total;
}

/*
 * NN-descent driver: float.rs:775-879 (sorted; also int8.rs:575-678,
 * uint8.rs:563-666) and float.rs:672-772 (unsorted, f32 only).
 * Output: raw heap distances, rows sorted ascending (deheap_sort).
 * *iters_done receives the number of completed iterations.
 */
int orc_nn_descent(int dtype, const void *x, int64_t n, int64_t d, int strict, int64_t k,
                   int64_t *rng_state, int64_t mc, int64_t n_iters, float delta,
                   float delta_improv, int use_delta_improv, const int32_t *leaf,
                   int64_t leaf_rows, int64_t leaf_cols, int use_sorted, int64_t T, int nw,
                   int32_t *out_idx, float *out_dist, int64_t *iters_done)
{
    orc_data D = {dtype, x, n, d, strict};
    orc_heap g;
    if (T < 1)
        T = 1;
    if (nw < 1)
        nw = 1;
    if (heap_make(&g, n, k))
        return -1;
    if (leaf && leaf_rows > 0)
        if (init_rp_tree(&D, &g, leaf, leaf_rows, leaf_cols, T, nw)) {
            heap_free(&g);
            return -1;
        }
    init_random(&D, &g, k, rng_state, nw);

    int64_t block = 65536 / T;
    int64_t n_blocks = n / block;
    const int64_t max_updates = (mc * mc + mc * (mc > 0 ? mc - 1 : 0) / 2) * block;
    int32_t *newc = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n * mc + 1));
    int32_t *oldc = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n * mc + 1));
    float *thr = (float *)malloc(sizeof(float) * (size_t)n);
    float *arr = (float *)malloc(sizeof(float) * (size_t)(T * max_updates * 3 + 3));
    int32_t *counts = (int32_t *)malloc(sizeof(int32_t) * (size_t)(T * (T + 1)));
    if (!newc || !oldc || !thr || !arr || !counts) {
        free(newc);
        free(oldc);
        free(thr);
        free(arr);
        free(counts);
        heap_free(&g);
        return -1;
    }
    orc_sorted_buf B = {arr, max_updates, counts};
    double prev = 0.0;
    int have_prev = 0;
    int64_t it;
    int rc = 0;
    for (it = 0; it < n_iters; it++) {
        if (build_candidates(&g, mc, rng_state, T, nw, newc, oldc)) {
            rc = -1;
            break;
        }
        uint64_t c = 0;
        for (int64_t b = 0; b <= n_blocks; b++) {
            int64_t b0 = b * block, b1 = (b + 1) * block < n ? (b + 1) * block : n;
            if (b0 >= b1)
                continue;
            for (int64_t v = 0; v < n; v++)
                thr[v] = g.dist[v * k];
            if (use_sorted) {
                gen_sorted_updates(&D, &B, newc + b0 * mc, oldc + b0 * mc, b1 - b0, mc, mc, thr,
                                   T, nw);
                c += apply_sorted_updates(&g, &B, T, nw);
            } else {
                memset(counts, 0, sizeof(int32_t) * (size_t)T);
                gen_unsorted_updates(&D, arr, max_updates, counts, newc + b0 * mc, oldc + b0 * mc,
                                     b1 - b0, mc, mc, thr, T, nw);
                c += apply_unsorted_updates(&g, arr, max_updates, counts, T, nw);
            }
        }
        /* float.rs:847: c <= (delta * k * n) as u32, in f32 */
        float lim = delta * (float)k * (float)n;
        uint32_t ulim = lim <= 0.0f ? 0u : (lim >= 4294967295.0f ? 4294967295u : (uint32_t)lim);
        if ((uint32_t)c <= ulim) {
            it++;
            break;
        }
        if (use_delta_improv) {
            double s = 0.0;
            for (int64_t e = 0; e < n * k; e++)
                if (g.dist[e] < ORC_INF)
                    s += (double)g.dist[e];
            if (have_prev) {
                double rel = fabs(s - prev) / fabs(prev);
                if (rel < (double)delta_improv) {
                    it++;
                    break;
                }
            }
            prev = s;
            have_prev = 1;
        }
        block = (block < n ? block : n);
        block = block * 2 > 1 ? block * 2 : 1;
        n_blocks = n / block;
    }
    if (iters_done)
        *iters_done = it;
    if (rc == 0) {
        memcpy(out_idx, g.idx, sizeof(int32_t) * (size_t)(n * k));
        memcpy(out_dist, g.dist, sizeof(float) * (size_t)(n * k));
        orc_deheap_sort(out_idx, out_dist, n, k, nw);
    }
    free(newc);
    free(oldc);
    free(thr);
    free(arr);
    free(counts);
    heap_free(&g);
    return rc;
}

/* ------------------------------------------------------------------------- */
/* Brute force (src/knn.rs:189-281): self skipped, strict '<', heap order.    */
/* ------------------------------------------------------------------------- */
void orc_brute_force_knn(int dtype, const void *x, int64_t n, int64_t d, int64_t n_neighbors,
                         int32_t *out_idx, float *out_dist, int nw)
{
    int64_t k = n_neighbors < (n > 0 ? n - 1 : 0) ? n_neighbors : (n > 0 ? n - 1 : 0);
    if (k < 1)
        k = 1;
    (void)nw;
#pragma omp parallel for num_threads(nw) schedule(dynamic, 16) if (nw > 1)
    for (int64_t i = 0; i < n; i++) {
        float *hd = out_dist + i * k;
        int32_t *hi = out_idx + i * k;
        for (int64_t t = 0; t < k; t++) {
            hd[t] = INFINITY;
            hi[t] = -1;
        }
        for (int64_t j = 0; j < n; j++) {
            if (i == j)
                continue;
            float dd;
            if (dtype == ORC_F32) {
                const float *a = (const float *)x + i * d, *b = (const float *)x + j * d;
                float dot = 0.0f;
                for (int64_t t = 0; t < d; t++)
                    dot += a[t] * b[t];
                dd = dot > 0.0f ? -dot : FLT_MIN;
            } else {
                orc_data D = {dtype, x, n, d, 0};
                dd = dist(&D, i, j);
            }
            if (dd < hd[0]) {
                hd[0] = dd;
                hi[0] = (int32_t)j;
                int64_t pos = 0;
                for (;;) {
                    int64_t l = pos * 2 + 1;
                    if (l >= k)
                        break;
                    int64_t r = l + 1, sw = pos;
                    if (hd[sw] < hd[l])
                        sw = l;
                    if (r < k && hd[sw] < hd[r])
                        sw = r;
                    if (sw == pos)
                        break;
                    float tf = hd[pos];
                    hd[pos] = hd[sw];
                    hd[sw] = tf;
                    int32_t ti = hi[pos];
                    hi[pos] = hi[sw];
                    hi[sw] = ti;
                    pos = sw;
                }
            }
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Exact k-NN with the NN-descent output layout (the contract of the GPU      */
/* exact mode): every j != i scored with the path's distance, plus self with  */
/* the seed value -1.0 (float.rs:210-214); k smallest by (distance, not-self, */
/* index); rows ascending.  acc: 0 = fast-math C dot (deterministic path),    */
/* 1 = serial f32 dot, 2 = f64 accumulation (f32 rows only).                  */
/* ------------------------------------------------------------------------- */
typedef struct {
    double d;
    int32_t self;
    int32_t j;
} orc_cand;

static inline int cand_less(const orc_cand *a, const orc_cand *b)
{
    if (a->d != b->d)
        return a->d < b->d;
    if (a->self != b->self)
        return a->self > b->self;
    return a->j < b->j;
}

void orc_exact_knn(int dtype, const void *x, int64_t n, int64_t d, int64_t k, int acc,
                   int include_self, int32_t *out_idx, double *out_dist, int nw)
{
    (void)nw;
#pragma omp parallel for num_threads(nw) schedule(dynamic, 8) if (nw > 1)
    for (int64_t i = 0; i < n; i++) {
        orc_cand *best = (orc_cand *)malloc(sizeof(orc_cand) * (size_t)(k + 1));
        int64_t m = 0;
        orc_data D = {dtype, x, n, d, acc == 0};
        for (int64_t j = 0; j < n; j++) {
            orc_cand c;
            c.j = (int32_t)j;
            c.self = (j == i);
            if (j == i) {
                if (!include_self)
                    continue;
                c.d = -1.0;
            } else if (dtype == ORC_F32 && acc == 2) {
                const float *a = (const float *)x + i * d, *b = (const float *)x + j * d;
                double s = 0.0;
                for (int64_t t = 0; t < d; t++)
                    s += (double)a[t] * (double)b[t];
                c.d = s > 0.0 ? -s : (double)FLT_MIN;
            } else {
                c.d = (double)dist(&D, i, j);
            }
            /* insertion into the sorted prefix best[0..m) */
            if (m == k && !cand_less(&c, &best[k - 1]))
                continue;
            int64_t pos = m < k ? m : k - 1;
            while (pos > 0 && cand_less(&c, &best[pos - 1])) {
                best[pos] = best[pos - 1];
                pos--;
            }
            best[pos] = c;
            if (m < k)
                m++;
        }
        for (int64_t t = 0; t < k; t++) {
            out_idx[i * k + t] = t < m ? best[t].j : -1;
            out_dist[i * k + t] = t < m ? best[t].d : (double)FLT_MAX;
        }
        free(best);
    }
}

/* ------------------------------------------------------------------------- */
/* Distance transforms: src/knn.rs:156-187                                    */
/* ------------------------------------------------------------------------- */
void orc_transform(int dtype, float *dd, int64_t m)
{
    for (int64_t i = 0; i < m; i++) {
        float v = dd[i];
        if (dtype == ORC_F32) {
            if (v >= 0.0f)
                dd[i] = 0.0f;
            else {
                double sim = (double)(-v);
                if (sim >= 1.0)
                    dd[i] = 0.0f;
                else {
                    double t = -log2(sim);
                    dd[i] = (float)(t > 0.0 ? t : 0.0);
                }
            }
        } else if (dtype == ORC_I8) {
            dd[i] = v >= 0.0f ? 0.0f : 1.0f / (-v);
        } else {
            dd[i] = v >= 0.0f ? 0.0f : (float)(-log2((double)(-v)));
        }
    }
}

/* ------------------------------------------------------------------------- */
/* src/knn.rs:78-113 prepare_float (in place on a caller-owned copy)          */
/* ------------------------------------------------------------------------- */
void orc_prepare_float(float *x, int64_t n, int64_t d)
{
    int all_unit = 1;
    for (int64_t i = 0; i < n && all_unit; i++) {
        float s = 0.0f;
        for (int64_t t = 0; t < d; t++)
            s += x[i * d + t] * x[i * d + t];
        s = sqrtf(s);
        if (fabsf(s - 1.0f) > 1e-5f)
            all_unit = 0;
    }
    if (all_unit)
        return;
    for (int64_t i = 0; i < n; i++) {
        float s = 0.0f;
        for (int64_t t = 0; t < d; t++)
            s += x[i * d + t] * x[i * d + t];
        s = sqrtf(s);
        if (s == 0.0f)
            s = 1.0f;
        for (int64_t t = 0; t < d; t++)
            x[i * d + t] /= s;
    }
}

/* ------------------------------------------------------------------------- */
/* knn_graph: src/knn.rs:345-454                                              */
/* ------------------------------------------------------------------------- */
typedef struct {
    int64_t n_neighbors;
    int64_t n_trees;        /* <= 0: clamp(T, 4, 8) */
    int64_t leaf_size;      /* <= 0: max(10, k) */
    int64_t max_candidates; /* <= 0: min(60, floor(1.5 k)) */
    int64_t max_rptree_depth;
    int64_t n_iters;        /* <= 0: max(5, round(log2 n)) */
    float delta;
    float delta_improv;
    int32_t use_sorted_updates;
    int32_t deterministic;
    int64_t brute_force_max_n;
} orc_knn_options;

#define ORC_ERR_EMPTY -2
#define ORC_ERR_NEIGHBORS -3
#define ORC_ERR_OOM -1

/*
 * `x` is modified in place by prepare_float for f32 input (pass a copy).
 * T = rayon::current_num_threads() of the caller (forced to 1 when
 * deterministic, src/knn.rs:350-355); nw = OS threads used to run the T workers.
 * Outputs are n x k_eff where k_eff = min(n_neighbors, n-1).max(1) (written to *k_out).
 */
int orc_knn_graph(int dtype, void *x, int64_t n, int64_t d, const orc_knn_options *o,
                  orc_nprng *rng, int64_t T, int nw, int32_t *out_idx, float *out_dist,
                  int64_t *k_out, int64_t *iters_done)
{
    if (o->n_neighbors <= 0)
        return ORC_ERR_NEIGHBORS;
    if (n <= 0)
        return ORC_ERR_EMPTY;
    if (o->deterministic)
        T = 1;
    if (T < 1)
        T = 1;
    if (dtype == ORC_F32)
        orc_prepare_float((float *)x, n, d);
    int64_t k = o->n_neighbors < n - 1 ? o->n_neighbors : n - 1;
    if (k < 1)
        k = 1;
    *k_out = k;
    int64_t n_trees = o->n_trees > 0 ? o->n_trees : (T < 4 ? 4 : (T > 8 ? 8 : T));
    int64_t leaf_size = o->leaf_size > 0 ? o->leaf_size : (k > 10 ? k : 10);
    int64_t n_iters = o->n_iters > 0 ? o->n_iters : (int64_t)round(log2((double)n));
    if (o->n_iters <= 0 && n_iters < 5)
        n_iters = 5;
    int64_t mc = o->max_candidates > 0 ? o->max_candidates : (int64_t)((float)k * 1.5f);
    if (o->max_candidates <= 0 && mc > 60)
        mc = 60;
    int64_t st[3];
    orc_nprng_randint3_for_tau(rng, st);
    if (n <= o->brute_force_max_n && dtype != ORC_F32) {
        orc_brute_force_knn(dtype, x, n, d, k, out_idx, out_dist, nw);
        if (iters_done)
            *iters_done = 0;
    } else {
        int64_t rows = 0, cols = 0;
        int32_t *leaf = orc_make_forest(dtype, x, n, d, k, n_trees, leaf_size, rng,
                                        o->max_rptree_depth, &rows, &cols, nw);
        if (!leaf && n_trees > 0)
            return ORC_ERR_OOM;
        int rc = orc_nn_descent(dtype, x, n, d, o->deterministic != 0, k, st, mc, n_iters,
                                o->delta, o->delta_improv, 1, leaf, rows, cols,
                                dtype == ORC_F32 ? o->use_sorted_updates : 1, T, nw, out_idx,
                                out_dist, iters_done);
        free(leaf);
        if (rc)
            return ORC_ERR_OOM;
    }
    orc_transform(dtype, out_dist, n * k);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Fuzzy simplicial set (the step after knn_graph): src/graph_construction.rs */
/* ------------------------------------------------------------------------- */
#define ORC_SMOOTH_K_TOLERANCE 1e-5f /* src/graph_construction.rs:12 */
#define ORC_MIN_K_DIST_SCALE 1e-3f   /* src/graph_construction.rs:13 */

/* smooth_knn_dist, src/graph_construction.rs:15-81.  Per row, in f64 and in row
 * order: rho = the first positive distance; sigma from at most 64 bisection
 * steps on sum_{j >= 1} exp(-(d_j - rho) / sigma) = log2(n_neighbors) (stop
 * when within 1e-5); floored at 1e-3 x the row mean (rho > 0) or 1e-3 x the
 * mean of all distances (rho = 0).  The reference takes that global mean with
 * ndarray's f32 mean; here it is an f64 sum rounded to f32 (it only enters the
 * floor of rows without a positive distance). */
void orc_smooth_knn_dist(const float *dist, int64_t n, int64_t k, float n_neighbors, float *sigma,
                         float *rho)
{
    const double target = log2((double)n_neighbors);
    double tot = 0.0;
    for (int64_t i = 0; i < n * k; i++) tot += (double)dist[i];
    const double mean_all = n * k > 0 ? (double)(float)(tot / (double)(n * k)) : 0.0;
    const double tol = (double)ORC_SMOOTH_K_TOLERANCE;
    for (int64_t i = 0; i < n; i++) {
        const float *row = dist + i * k;
        double rho_i = 0.0;
        for (int64_t j = 0; j < k; j++)
            if ((double)row[j] > 0.0) {
                rho_i = (double)row[j];
                break;
            }
        double lo = 0.0, hi = INFINITY, mid = 1.0;
        for (int it = 0; it < 64; it++) {
            double psum = 0.0;
            for (int64_t j = 1; j < k; j++) {
                const double d = (double)row[j] - rho_i;
                psum += d > 0.0 ? exp(-(d / mid)) : 1.0;
            }
            if (fabs(psum - target) < tol) break;
            if (psum > target) {
                hi = mid;
                mid = (lo + hi) / 2.0;
            } else {
                lo = mid;
                mid = isinf(hi) ? mid * 2.0 : (lo + hi) / 2.0;
            }
        }
        double s = mid;
        if (rho_i > 0.0) {
            double m = 0.0;
            for (int64_t j = 0; j < k; j++) m += (double)row[j];
            m /= (double)k;
            const double fl = (double)ORC_MIN_K_DIST_SCALE * m;
            if (s < fl) s = fl;
        } else {
            const double fl = (double)ORC_MIN_K_DIST_SCALE * mean_all;
            if (s < fl) s = fl;
        }
        sigma[i] = (float)s;
        rho[i] = (float)rho_i;
    }
}

/* compute_membership_strengths, src/graph_construction.rs:83-115 (f32):
 * self -> 0, d - rho <= 0 or sigma == 0 -> 1, else exp(-(d - rho) / sigma). */
static float membership(int64_t i, int32_t idx, float d, float sg, float rh)
{
    if (idx == (int32_t)i) return 0.0f;
    if (d - rh <= 0.0f || sg == 0.0f) return 1.0f;
    return expf(-((d - rh) / sg));
}

typedef struct {
    int64_t n;
    int64_t *ptr;
    int32_t *col;
    float *val;
} orc_csr;

static int cmp_col_val(const void *a, const void *b)
{
    const int32_t ca = *(const int32_t *)a, cb = *(const int32_t *)b;
    return (ca > cb) - (ca < cb);
}

/* COO triplets (row-major by construction) -> CSR with each row sorted by column;
 * duplicates (none in a kNN graph) would be summed, as TriMat::to_csr does. */
static int coo_to_csr(int64_t n, int64_t nnz, const int32_t *r, const int32_t *c, const float *v,
                      orc_csr *out)
{
    out->n = n;
    out->ptr = calloc((size_t)n + 1, sizeof(int64_t));
    out->col = malloc(((size_t)nnz + 1) * sizeof(int32_t));
    out->val = malloc(((size_t)nnz + 1) * sizeof(float));
    typedef struct { int32_t c; float v; } cv;
    cv *tmp = malloc(((size_t)nnz + 1) * sizeof(cv));
    int64_t *fill = calloc((size_t)n + 1, sizeof(int64_t));
    if (!out->ptr || !out->col || !out->val || !tmp || !fill) return -1;
    for (int64_t e = 0; e < nnz; e++) out->ptr[r[e] + 1]++;
    for (int64_t i = 0; i < n; i++) out->ptr[i + 1] += out->ptr[i];
    for (int64_t e = 0; e < nnz; e++) {
        const int64_t p = out->ptr[r[e]] + fill[r[e]]++;
        tmp[p].c = c[e];
        tmp[p].v = v[e];
    }
    int64_t w = 0;
    for (int64_t i = 0; i < n; i++) {
        const int64_t a = out->ptr[i], b = out->ptr[i + 1];
        qsort(tmp + a, (size_t)(b - a), sizeof(cv), cmp_col_val);
        const int64_t start = w;
        for (int64_t p = a; p < b; p++) {
            if (w > start && out->col[w - 1] == tmp[p].c) {
                out->val[w - 1] += tmp[p].v;
            } else {
                out->col[w] = tmp[p].c;
                out->val[w] = tmp[p].v;
                w++;
            }
        }
        out->ptr[i] = start;
    }
    out->ptr[n] = w;
    free(tmp);
    free(fill);
    return 0;
}

static void csr_free(orc_csr *m)
{
    free(m->ptr);
    free(m->col);
    free(m->val);
}

/* Row-wise sparse binary op over sorted rows, dropping exact zeros
 * (scipy_csr_rowwise_binop, src/csr_matmul.rs:81-147): op 0 = a + b (union),
 * 1 = a * b (intersection), 2 = a - b (union); a missing operand is 0.0. */
static int csr_binop(const orc_csr *A, const orc_csr *B, int op, orc_csr *C)
{
    const int64_t n = A->n;
    const int64_t cap = A->ptr[n] + B->ptr[n] + 1;
    C->n = n;
    C->ptr = calloc((size_t)n + 1, sizeof(int64_t));
    C->col = malloc((size_t)cap * sizeof(int32_t));
    C->val = malloc((size_t)cap * sizeof(float));
    if (!C->ptr || !C->col || !C->val) return -1;
    int64_t w = 0;
    for (int64_t i = 0; i < n; i++) {
        int64_t ja = A->ptr[i], jb = B->ptr[i];
        const int64_t ea = A->ptr[i + 1], eb = B->ptr[i + 1];
        while (ja < ea || jb < eb) {
            const int64_t ca = ja < ea ? A->col[ja] : INT64_MAX;
            const int64_t cb = jb < eb ? B->col[jb] : INT64_MAX;
            const int64_t c = ca < cb ? ca : cb;
            const int has_a = ca == c, has_b = cb == c;
            const float va = has_a ? A->val[ja] : 0.0f, vb = has_b ? B->val[jb] : 0.0f;
            int emit = op == 1 ? (has_a && has_b) : 1;
            float v = op == 0 ? va + vb : (op == 1 ? va * vb : va - vb);
            if (emit && v != 0.0f) {
                C->col[w] = (int32_t)c;
                C->val[w] = v;
                w++;
            }
            if (has_a) ja++;
            if (has_b) jb++;
        }
        C->ptr[i + 1] = w;
    }
    return 0;
}

static int csr_transpose(const orc_csr *A, orc_csr *T)
{
    const int64_t n = A->n, nnz = A->ptr[n];
    int32_t *r = malloc(((size_t)nnz + 1) * sizeof(int32_t));
    int32_t *c = malloc(((size_t)nnz + 1) * sizeof(int32_t));
    if (!r || !c) return -1;
    for (int64_t i = 0; i < n; i++)
        for (int64_t p = A->ptr[i]; p < A->ptr[i + 1]; p++) {
            r[p] = A->col[p];
            c[p] = (int32_t)i;
        }
    const int rc = coo_to_csr(n, nnz, r, c, A->val, T);
    free(r);
    free(c);
    return rc;
}

/* neighbor_graph_matrix_with_coo, src/graph_construction.rs:175-226 (without
 * the reference-value override, which needs on-disk goldens): membership
 * strengths of the kNN graph (nonzero only), then with `symmetrize` the fuzzy
 * union A + A^T - A o A^T, rows sorted by column, zeros dropped
 * (canonicalize_csr_rows, :228-240).  Writes sigma/rho (n), indptr (n + 1),
 * indices / data (capacity cap) and returns nnz, or -1 (allocation, capacity). */
int64_t orc_fuzzy_graph(const int32_t *inds, const float *dist, int64_t n, int64_t k, float n_neighbors,
                        int symmetrize, float *sigma, float *rho, int64_t *indptr, int32_t *indices,
                        float *data, int64_t cap)
{
    orc_smooth_knn_dist(dist, n, k, n_neighbors, sigma, rho);
    int32_t *r = malloc(((size_t)(n * k) + 1) * sizeof(int32_t));
    int32_t *c = malloc(((size_t)(n * k) + 1) * sizeof(int32_t));
    float *v = malloc(((size_t)(n * k) + 1) * sizeof(float));
    if (!r || !c || !v) return -1;
    int64_t m = 0;
    for (int64_t i = 0; i < n; i++)
        for (int64_t j = 0; j < k; j++) {
            const int32_t idx = inds[i * k + j];
            if (idx == -1) continue;
            const float val = membership(i, idx, dist[i * k + j], sigma[i], rho[i]);
            if (val != 0.0f) {
                r[m] = (int32_t)i;
                c[m] = idx;
                v[m] = val;
                m++;
            }
        }
    orc_csr A, res;
    if (coo_to_csr(n, m, r, c, v, &A)) return -1;
    free(r);
    free(c);
    free(v);
    if (symmetrize) {
        orc_csr T, sum, prod;
        if (csr_transpose(&A, &T) || csr_binop(&A, &T, 1, &prod) || csr_binop(&A, &T, 0, &sum) ||
            csr_binop(&sum, &prod, 2, &res))
            return -1;
        csr_free(&T);
        csr_free(&sum);
        csr_free(&prod);
        csr_free(&A);
    } else {
        res = A;
    }
    const int64_t nnz = res.ptr[n];
    int64_t w = 0;
    if (nnz <= cap) {
        for (int64_t i = 0; i < n; i++) {
            indptr[i] = w;
            for (int64_t p = res.ptr[i]; p < res.ptr[i + 1]; p++)
                if (res.val[p] != 0.0f) {
                    indices[w] = res.col[p];
                    data[w] = res.val[p];
                    w++;
                }
        }
        indptr[n] = w;
    }
    csr_free(&res);
    return nnz <= cap ? w : -1;
}
