/*
 * oracle.c -- CPU oracle for the SPID-Join equi-join path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the reported CPU baseline).  The
 * product library never links or calls it.  Every function cites the
 * reference file:line it restates; the header lists what pins each one.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

/* ------------------------------------------------------------------------- */
/* 1. Hashes                                                                 */
/* ------------------------------------------------------------------------- */

static uint32_t g_crc[256];
static pthread_once_t g_crc_once = PTHREAD_ONCE_INIT;

/* The reflected CRC-32 table (polynomial 0xEDB88320), generated from the
 * polynomial rather than copied: reference src/dpu/src/hash.c:15-59 holds the
 * same 256 words as a literal. */
static void crc_init(void)
{
    for (uint32_t n = 0; n < 256; n++) {
        uint32_t c = n;
        for (int k = 0; k < 8; k++) c = (c & 1u) ? (0xEDB88320u ^ (c >> 1)) : (c >> 1);
        g_crc[n] = c;
    }
}

const uint32_t *or_crc_table(void)
{
    pthread_once(&g_crc_once, crc_init);
    return g_crc;
}

/* Paul Hsieh's SuperFastHash over the 4 little-endian bytes of the key, with
 * the initial hash equal to the length (4): one main-loop round, no tail,
 * then the published avalanche.  Reference: hash.c:117-166. */
uint32_t or_glb_partition_hash(uint32_t key)
{
    uint32_t h = 4u;
    h += key & 0xFFFFu;
    uint32_t tmp = ((key >> 16) << 11) ^ h;
    h = (h << 16) ^ tmp;
    h += h >> 11;
    h ^= h << 3;
    h += h >> 5;
    h ^= h << 4;
    h += h >> 17;
    h ^= h << 25;
    h += h >> 6;
    return h;
}

/* Tabulation over the CRC table: (M & T[b0]) ^ T[b1] ^ T[b2] ^ T[b3]; '&'
 * binds tighter than '^'.  Reference: join_hash hash.c:72-83 (M = 0xd9d65adc,
 * bytes of key) and local_partition_hash hash.c:168-178 (M = 0xbfd06116,
 * bytes of key >> 6). */
static inline uint32_t tab_hash(uint32_t mask, uint32_t k)
{
    const uint32_t *T = or_crc_table();
    return (mask & T[k & 0xFFu]) ^ T[(k >> 8) & 0xFFu] ^ T[(k >> 16) & 0xFFu] ^ T[(k >> 24) & 0xFFu];
}

uint32_t or_join_hash(uint32_t key) { return tab_hash(0xd9d65adcu, key); }
uint32_t or_local_partition_hash(uint32_t key) { return tab_hash(0xbfd06116u, key >> 6); }

/* Bob Jenkins' one-at-a-time hash over the 4 little-endian key bytes.
 * Reference: hash.c:85-112. */
uint32_t or_join_hash2(uint32_t key)
{
    uint32_t h = 0;
    for (int i = 0; i < 4; i++) {
        h += (key >> (8 * i)) & 0xFFu;
        h += h << 10;
        h ^= h >> 6;
    }
    h += h << 3;
    h ^= h >> 11;
    h += h << 15;
    return h;
}

void or_hash_batch(int which, const uint32_t *keys, uint32_t *out, uint64_t n)
{
    for (uint64_t i = 0; i < n; i++) {
        uint32_t k = keys[i];
        switch (which) {
        case 0: out[i] = or_glb_partition_hash(k); break;
        case 1: out[i] = or_join_hash(k); break;
        case 2: out[i] = or_local_partition_hash(k); break;
        default: out[i] = or_join_hash2(k); break;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Small parallel-for helper                                                 */
/* ------------------------------------------------------------------------- */

typedef void (*pf_fn)(void *ctx, uint64_t item, int tid);
typedef struct {
    pf_fn fn;
    void *ctx;
    uint64_t n;
    _Atomic uint64_t next;
    int tid_base;
} pf_state_t;
typedef struct { pf_state_t *st; int tid; } pf_arg_t;

static void *pf_worker(void *p)
{
    pf_arg_t *a = (pf_arg_t *)p;
    for (;;) {
        uint64_t i = atomic_fetch_add(&a->st->next, 1);
        if (i >= a->st->n) break;
        a->st->fn(a->st->ctx, i, a->tid);
    }
    return NULL;
}

static int num_threads(int req)
{
    if (req > 0) return req;
    long c = sysconf(_SC_NPROCESSORS_ONLN);
    return c > 0 ? (int)c : 1;
}

static void parallel_for(int threads, uint64_t n, pf_fn fn, void *ctx)
{
    pf_state_t st;
    st.fn = fn;
    st.ctx = ctx;
    st.n = n;
    atomic_init(&st.next, 0);
    if (threads <= 1 || n <= 1) {
        for (uint64_t i = 0; i < n; i++) fn(ctx, i, 0);
        return;
    }
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)threads);
    pf_arg_t *args = (pf_arg_t *)malloc(sizeof(pf_arg_t) * (size_t)threads);
    for (int t = 0; t < threads; t++) {
        args[t].st = &st;
        args[t].tid = t;
        pthread_create(&th[t], NULL, pf_worker, &args[t]);
    }
    for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
    free(th);
    free(args);
}

/* ------------------------------------------------------------------------- */
/* 2. Closed-form oracle                                                     */
/* ------------------------------------------------------------------------- */

static inline uint64_t mix64(uint64_t x)
{
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static inline uint64_t pair_word(uint32_t r, uint32_t s) { return ((uint64_t)r << 32) | s; }

/* LSD radix sort of 64-bit words, 16 bits per pass. */
static void radix_sort_u64(uint64_t *a, uint64_t n)
{
    if (n < 2) return;
    uint64_t *tmp = (uint64_t *)malloc(sizeof(uint64_t) * n);
    if (!tmp) return;
    uint64_t *src = a, *dst = tmp;
    static const int B = 16;
    uint64_t *cnt = (uint64_t *)malloc(sizeof(uint64_t) << B);
    for (int pass = 0; pass < 4; pass++) {
        int sh = pass * B;
        memset(cnt, 0, sizeof(uint64_t) << B);
        for (uint64_t i = 0; i < n; i++) cnt[(src[i] >> sh) & 0xFFFFu]++;
        uint64_t run = 0;
        for (uint64_t d = 0; d < (1u << B); d++) { uint64_t c = cnt[d]; cnt[d] = run; run += c; }
        for (uint64_t i = 0; i < n; i++) dst[cnt[(src[i] >> sh) & 0xFFFFu]++] = src[i];
        uint64_t *t = src; src = dst; dst = t;
    }
    /* 4 passes: result is back in a */
    free(cnt);
    free(tmp);
}

void or_sort_pairs(or_pair_t *pairs, uint64_t n)
{
    uint64_t *w = (uint64_t *)malloc(sizeof(uint64_t) * (n ? n : 1));
    if (!w) return;
    for (uint64_t i = 0; i < n; i++) w[i] = pair_word(pairs[i].r_rid, pairs[i].s_rid);
    radix_sort_u64(w, n);
    for (uint64_t i = 0; i < n; i++) { pairs[i].r_rid = (uint32_t)(w[i] >> 32); pairs[i].s_rid = (uint32_t)w[i]; }
    free(w);
}

/* R sorted by key as (key << 32 | rid) words; zero keys dropped. */
static uint64_t *sorted_r(const or_tuple_t *R, uint64_t nR, uint64_t *n_out)
{
    uint64_t *w = (uint64_t *)malloc(sizeof(uint64_t) * (nR ? nR : 1));
    if (!w) return NULL;
    uint64_t m = 0;
    for (uint64_t i = 0; i < nR; i++)
        if (R[i].key != 0) w[m++] = ((uint64_t)R[i].key << 32) | R[i].rid;
    radix_sort_u64(w, m);
    *n_out = m;
    return w;
}

/* lower bound of key in the sorted (key<<32|rid) words */
static uint64_t lower_key(const uint64_t *w, uint64_t m, uint32_t key)
{
    uint64_t lo = 0, hi = m;
    while (lo < hi) {
        uint64_t mid = (lo + hi) >> 1;
        if ((uint32_t)(w[mid] >> 32) < key) lo = mid + 1; else hi = mid;
    }
    return lo;
}

/* Expected output of the PK-FK equi-join: for every S tuple with a non-zero
 * key, one pair (R rid, S rid) per R tuple of equal key (exactly one when R
 * keys are unique).  The probe that emits one pair per matching S tuple is
 * a2a_srj_probe_inner_mt.c:136-148. */
int64_t or_join_expected(const or_tuple_t *R, uint64_t nR, const or_tuple_t *S, uint64_t nS,
                         or_pair_t *out, uint64_t out_cap)
{
    uint64_t m;
    uint64_t *w = sorted_r(R, nR, &m);
    if (!w) return -2;
    uint64_t cnt = 0;
    for (uint64_t j = 0; j < nS; j++) {
        uint32_t k = S[j].key;
        if (k == 0) continue;
        for (uint64_t i = lower_key(w, m, k); i < m && (uint32_t)(w[i] >> 32) == k; i++) {
            if (cnt >= out_cap) { free(w); return -1; }
            out[cnt].r_rid = (uint32_t)w[i];
            out[cnt].s_rid = S[j].rid;
            cnt++;
        }
    }
    free(w);
    or_sort_pairs(out, cnt);
    return (int64_t)cnt;
}

void or_pairs_checksum(const or_pair_t *pairs, uint64_t n, uint64_t *sum, uint64_t *xr)
{
    uint64_t s = 0, x = 0;
    for (uint64_t i = 0; i < n; i++) {
        uint64_t h = mix64(pair_word(pairs[i].r_rid, pairs[i].s_rid));
        s += h;
        x ^= h;
    }
    *sum = s;
    *xr = x;
}

typedef struct {
    const or_tuple_t *S;
    uint64_t nS;
    const uint64_t *w; uint64_t m;          /* sparse path */
    const uint32_t *inv; const uint8_t *has; uint64_t maxkey; /* dense path */
    uint64_t chunk;
    uint64_t *cnt, *sum, *xr;               /* per chunk */
} cks_ctx_t;

static void cks_chunk(void *p, uint64_t c, int tid)
{
    (void)tid;
    cks_ctx_t *x = (cks_ctx_t *)p;
    uint64_t b = c * x->chunk, e = b + x->chunk;
    if (e > x->nS) e = x->nS;
    uint64_t cnt = 0, s = 0, xo = 0;
    for (uint64_t j = b; j < e; j++) {
        uint32_t k = x->S[j].key;
        if (k == 0) continue;
        if (x->inv) {
            if (k <= x->maxkey && x->has[k]) {
                uint64_t h = mix64(pair_word(x->inv[k], x->S[j].rid));
                cnt++; s += h; xo ^= h;
            }
        } else {
            for (uint64_t i = lower_key(x->w, x->m, k); i < x->m && (uint32_t)(x->w[i] >> 32) == k; i++) {
                uint64_t h = mix64(pair_word((uint32_t)x->w[i], x->S[j].rid));
                cnt++; s += h; xo ^= h;
            }
        }
    }
    x->cnt[c] = cnt; x->sum[c] = s; x->xr[c] = xo;
}

int or_join_checksum(const or_tuple_t *R, uint64_t nR, const or_tuple_t *S, uint64_t nS,
                     int threads, uint64_t *count, uint64_t *sum, uint64_t *xr)
{
    cks_ctx_t x;
    memset(&x, 0, sizeof(x));
    x.S = S;
    x.nS = nS;
    uint64_t maxkey = 0;
    for (uint64_t i = 0; i < nR; i++) if (R[i].key > maxkey) maxkey = R[i].key;
    uint32_t *inv = NULL;
    uint8_t *has = NULL;
    uint64_t *w = NULL;
    int dense_ok = maxkey <= (uint64_t)4 * nR + 1024 && maxkey < (1ull << 31);
    if (dense_ok) {
        /* Keys unique in R: inv[key] = rid (the inverse-permutation closed form). */
        inv = (uint32_t *)malloc(sizeof(uint32_t) * (maxkey + 1));
        has = (uint8_t *)calloc(maxkey + 1, 1);
        if (!inv || !has) { free(inv); free(has); return -2; }
        for (uint64_t i = 0; i < nR; i++) {
            uint32_t k = R[i].key;
            if (k == 0) continue;
            if (has[k]) { dense_ok = 0; break; }     /* duplicate keys: use the general path */
            has[k] = 1;
            inv[k] = R[i].rid;
        }
        if (!dense_ok) { free(inv); free(has); inv = NULL; has = NULL; }
    }
    if (!dense_ok) {
        uint64_t m;
        w = sorted_r(R, nR, &m);
        if (!w) return -2;
        x.w = w; x.m = m;
    } else {
        x.inv = inv; x.has = has; x.maxkey = maxkey;
    }
    x.chunk = 1u << 16;
    uint64_t nchunks = (nS + x.chunk - 1) / x.chunk;
    x.cnt = (uint64_t *)calloc(nchunks + 1, sizeof(uint64_t));
    x.sum = (uint64_t *)calloc(nchunks + 1, sizeof(uint64_t));
    x.xr = (uint64_t *)calloc(nchunks + 1, sizeof(uint64_t));
    parallel_for(num_threads(threads), nchunks, cks_chunk, &x);
    uint64_t c = 0, s = 0, xo = 0;
    for (uint64_t i = 0; i < nchunks; i++) { c += x.cnt[i]; s += x.sum[i]; xo ^= x.xr[i]; }
    *count = c; *sum = s; *xr = xo;
    free(x.cnt); free(x.sum); free(x.xr);
    free(inv); free(has); free(w);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* 3. CPU restatement of the reference join                                  */
/* ------------------------------------------------------------------------- */

#define NUM_DPU_RANK 64       /* type_defs.hpp:6 */
#define ELEM_PER_BLOCK2 256   /* glb_partition_count.c:36-38: 2048-byte read blocks */
#define FILL_BYTE 1935        /* a2a_packetwise_srj_local_partition.c:21 */
#define HT_SLOTS 4096         /* a2a_srj_build.c: 32 KB table of 8-byte slots */

static int ilog2u(uint32_t x) { int r = 0; while (x >>= 1) r++; return r; }

/* Tuples per virtual DPU under LoadColumn's split: the byte size is padded
 * up to a multiple of D tuples (a whole extra block when already aligned)
 * and cut into D equal contiguous chunks.  Reference: query.cc:239-250. */
static uint64_t tuples_per_dpu(uint64_t n, uint64_t D)
{
    uint64_t bytes = n * 8u, align = D * 8u;
    uint64_t total = bytes + (align - bytes % align);
    return total / 8u / D;
}

/* Within-rank DPU of replica b of a bank set: the LUT_BG8/16/32/64 tables of
 * glb_chipwise_partition_count.c:131-140 are (b / (B/8)) * 8 + b % (B/8). */
static uint32_t lut_bg(uint32_t b, uint32_t B)
{
    uint32_t w = B / 8u;
    return (b / w) * 8u + (b % w);
}

typedef struct {
    const or_tuple_t *R, *S;
    uint64_t nR, nS;
    uint32_t D, G, Rg, B;
    uint32_t nRpart;             /* number of distinct R partitions */
    uint64_t tpdR, tpdS;         /* tuples per source DPU */
    uint32_t *dpu_rpart;         /* D: which R partition a DPU holds */
    /* routing */
    uint32_t *cntR;              /* [srcR][nRpart] */
    uint32_t *cntS;              /* [srcS][D] */
    uint64_t *offR;              /* [srcR][nRpart] */
    uint64_t *offS;              /* [srcS][D] */
    or_tuple_t *Rp;              /* R grouped by partition */
    or_tuple_t *Sd;              /* S grouped by destination DPU */
    uint64_t *rstart;            /* nRpart+1 */
    uint64_t *sstart;            /* D+1 */
    int alg;
    or_pair_t *res;              /* results at S positions */
    uint64_t *rescnt;            /* per DPU */
    _Atomic int err;
} spj_ctx_t;

/* Destination partition of an R tuple.  Plain: SFH & (Rg*64-1)
 * (glb_partition_count.c:192-195 with P = Rg*64 from queryDpuFunc.cc:274-286).
 * SPID: SFH & (Rg*64/B - 1) (bg_broadcast_count.c:192, P from
 * queryDpuFunc.cc:496). */
static inline uint32_t r_partition(const spj_ctx_t *c, uint32_t key)
{
    return or_glb_partition_hash(key) & (c->nRpart - 1u);
}

/* Destination DPU of the S tuple at index i of its 256-tuple read block.
 * Plain: (SFH & (Rg*64-1)) + (i & (G-1)) * Rg*64 (glb_partition_count.c:315-318).
 * SPID: the chipwise mapping of glb_chipwise_partition_count.c:251-255. */
static inline uint32_t s_dest(const spj_ctx_t *c, uint32_t key, uint32_t i)
{
    uint32_t h = or_glb_partition_hash(key);
    uint32_t group_dpus = c->Rg * NUM_DPU_RANK;
    if (c->B == 1) return (h & (group_dpus - 1u)) + (i & (c->G - 1u)) * group_dpus;
    uint32_t q = h & (c->nRpart - 1u);
    uint32_t bgpr = NUM_DPU_RANK / c->B;                 /* bank sets per rank */
    uint32_t shiftval = (uint32_t)ilog2u(c->B) - 3u;
    uint32_t slot = ELEM_PER_BLOCK2 / c->G;
    uint32_t d = ((q & (bgpr - 1u)) << shiftval) + ((q >> ilog2u(bgpr)) << 6) + lut_bg(i & (c->B - 1u), c->B);
    d += ((i / slot) & (c->G - 1u)) << ilog2u(group_dpus);
    return d;
}

static void route_count(void *p, uint64_t src, int tid)
{
    (void)tid;
    spj_ctx_t *c = (spj_ctx_t *)p;
    if (src < c->D) {           /* R source chunk */
        uint32_t *cnt = c->cntR + src * c->nRpart;
        uint64_t b = src * c->tpdR, e = b + c->tpdR;
        if (e > c->nR) e = c->nR;
        for (uint64_t j = b; j < e; j++)
            if (c->R[j].key) cnt[r_partition(c, c->R[j].key)]++;
    } else {                    /* S source chunk */
        src -= c->D;
        uint32_t *cnt = c->cntS + src * c->D;
        uint64_t b = src * c->tpdS, e = b + c->tpdS;
        if (e > c->nS) e = c->nS;
        for (uint64_t j = b; j < e; j++)
            if (c->S[j].key) cnt[s_dest(c, c->S[j].key, (uint32_t)((j - b) % ELEM_PER_BLOCK2))]++;
    }
}

static void route_scatter(void *p, uint64_t src, int tid)
{
    (void)tid;
    spj_ctx_t *c = (spj_ctx_t *)p;
    if (src < c->D) {
        uint64_t *off = c->offR + src * c->nRpart;
        uint64_t b = src * c->tpdR, e = b + c->tpdR;
        if (e > c->nR) e = c->nR;
        for (uint64_t j = b; j < e; j++)
            if (c->R[j].key) c->Rp[off[r_partition(c, c->R[j].key)]++] = c->R[j];
    } else {
        src -= c->D;
        uint64_t *off = c->offS + src * c->D;
        uint64_t b = src * c->tpdS, e = b + c->tpdS;
        if (e > c->nS) e = c->nS;
        for (uint64_t j = b; j < e; j++)
            if (c->S[j].key) c->Sd[off[s_dest(c, c->S[j].key, (uint32_t)((j - b) % ELEM_PER_BLOCK2))]++] = c->S[j];
    }
}

/* P_loc = 1 << bitlen(n / 1935): a2a_packetwise_srj_local_partition.c:323-337. */
static uint32_t local_partitions(uint64_t n)
{
    uint64_t t = n / FILL_BYTE;
    uint32_t p = 1;
    while (t > 0) { t >>= 1; p <<= 1; }
    return p;
}

static int cmp_tuple_key(const void *a, const void *b)
{
    uint32_t x = ((const or_tuple_t *)a)->key, y = ((const or_tuple_t *)b)->key;
    return (x > y) - (x < y);
}

/* One virtual DPU: local partition, then per local partition build + probe
 * (phj), or sort + merge (smj), or nested loop (nl). */
static void dpu_join(void *p, uint64_t d, int tid)
{
    (void)tid;
    spj_ctx_t *c = (spj_ctx_t *)p;
    uint32_t rp = c->dpu_rpart[d];
    const or_tuple_t *Rd = c->Rp + c->rstart[rp];
    uint64_t nr = c->rstart[rp + 1] - c->rstart[rp];
    const or_tuple_t *Sd = c->Sd + c->sstart[d];
    uint64_t ns = c->sstart[d + 1] - c->sstart[d];
    or_pair_t *out = c->res + c->sstart[d];
    uint64_t cnt = 0;

    if (c->alg == 3) {
        /* nested_loop_join.c:184-231: first equal R key wins */
        for (uint64_t j = 0; j < ns; j++)
            for (uint64_t i = 0; i < nr; i++)
                if (Rd[i].key == Sd[j].key) { out[cnt].r_rid = Rd[i].rid; out[cnt].s_rid = Sd[j].rid; cnt++; break; }
        c->rescnt[d] = cnt;
        return;
    }
    if (c->alg == 2) {
        /* mpsm_join_sort_probe.c:152-236: sort both sides by key, then one merge
         * pass; an S tuple takes the rid of the equal R key.  (The reference
         * also range-buckets by key first, mpsm_join_local_partition.c; the
         * buckets only bound the WRAM working set.) */
        or_tuple_t *rs = (or_tuple_t *)malloc(sizeof(or_tuple_t) * (nr ? nr : 1));
        or_tuple_t *ss = (or_tuple_t *)malloc(sizeof(or_tuple_t) * (ns ? ns : 1));
        if (!rs || !ss) { free(rs); free(ss); atomic_store(&c->err, -2); return; }
        memcpy(rs, Rd, sizeof(or_tuple_t) * nr);
        memcpy(ss, Sd, sizeof(or_tuple_t) * ns);
        qsort(rs, nr, sizeof(or_tuple_t), cmp_tuple_key);
        qsort(ss, ns, sizeof(or_tuple_t), cmp_tuple_key);
        uint64_t i = 0;
        for (uint64_t j = 0; j < ns; j++) {
            while (i < nr && rs[i].key < ss[j].key) i++;
            if (i < nr && rs[i].key == ss[j].key) { out[cnt].r_rid = rs[i].rid; out[cnt].s_rid = ss[j].rid; cnt++; }
        }
        free(rs);
        free(ss);
        c->rescnt[d] = cnt;
        return;
    }

    /* phj: local partition (hist + prefix + scatter) of R, S with R's P_loc */
    uint32_t P = local_partitions(nr);
    uint64_t *hr = (uint64_t *)calloc((size_t)P + 1, sizeof(uint64_t));
    uint64_t *hs = (uint64_t *)calloc((size_t)P + 1, sizeof(uint64_t));
    or_tuple_t *lr = (or_tuple_t *)malloc(sizeof(or_tuple_t) * (nr ? nr : 1));
    or_tuple_t *ls = (or_tuple_t *)malloc(sizeof(or_tuple_t) * (ns ? ns : 1));
    uint32_t *tk = (uint32_t *)malloc(sizeof(uint32_t) * HT_SLOTS);
    uint32_t *tv = (uint32_t *)malloc(sizeof(uint32_t) * HT_SLOTS);
    if (!hr || !hs || !lr || !ls || !tk || !tv) {
        atomic_store(&c->err, -2);
        goto done;
    }
    for (uint64_t i = 0; i < nr; i++) hr[(or_local_partition_hash(Rd[i].key) & (P - 1u)) + 1]++;
    for (uint64_t j = 0; j < ns; j++) hs[(or_local_partition_hash(Sd[j].key) & (P - 1u)) + 1]++;
    for (uint32_t q = 0; q < P; q++) { hr[q + 1] += hr[q]; hs[q + 1] += hs[q]; }
    {
        uint64_t *wr = (uint64_t *)malloc(sizeof(uint64_t) * P);
        uint64_t *ws = (uint64_t *)malloc(sizeof(uint64_t) * P);
        memcpy(wr, hr, sizeof(uint64_t) * P);
        memcpy(ws, hs, sizeof(uint64_t) * P);
        for (uint64_t i = 0; i < nr; i++) lr[wr[or_local_partition_hash(Rd[i].key) & (P - 1u)]++] = Rd[i];
        for (uint64_t j = 0; j < ns; j++) ls[ws[or_local_partition_hash(Sd[j].key) & (P - 1u)]++] = Sd[j];
        free(wr);
        free(ws);
    }
    for (uint32_t q = 0; q < P; q++) {
        uint64_t rn = hr[q + 1] - hr[q];
        if (rn > HT_SLOTS) { atomic_store(&c->err, -3); goto done; }
        /* build: a2a_srj_build.c:202-296 -- slot = join_hash & 0xFFF, linear
         * probing with wrap, key 0 = empty */
        memset(tk, 0, sizeof(uint32_t) * HT_SLOTS);
        for (uint64_t i = hr[q]; i < hr[q + 1]; i++) {
            uint32_t k = lr[i].key, h = or_join_hash(k) & (HT_SLOTS - 1u);
            while (tk[h] != 0) {
                if (tk[h] == k) { atomic_store(&c->err, -4); goto done; }
                h = (h + 1u) & (HT_SLOTS - 1u);
            }
            tk[h] = k;
            tv[h] = lr[i].rid;
        }
        /* probe: a2a_srj_probe_inner_mt.c:118-168 -- stop at empty, first
         * match, or after 4097 steps */
        for (uint64_t j = hs[q]; j < hs[q + 1]; j++) {
            uint32_t k = ls[j].key, h = or_join_hash(k) & (HT_SLOTS - 1u);
            for (int runs = 0; runs <= HT_SLOTS; runs++) {
                if (tk[h] == 0) break;
                if (tk[h] == k) { out[cnt].r_rid = tv[h]; out[cnt].s_rid = ls[j].rid; cnt++; break; }
                h = (h + 1u) & (HT_SLOTS - 1u);
            }
        }
    }
done:
    c->rescnt[d] = cnt;
    free(hr); free(hs); free(lr); free(ls); free(tk); free(tv);
}

int64_t or_spid_join_cpu(const or_tuple_t *R, uint64_t nR, const or_tuple_t *S, uint64_t nS,
                         const or_design_t *design, or_pair_t *out, uint64_t out_cap)
{
    spj_ctx_t c;
    memset(&c, 0, sizeof(c));
    int ranks = design->rank_num, G = design->rank_set, B = design->bank_set;
    if (ranks <= 0 || (ranks & (ranks - 1)) || G <= 0 || (G & (G - 1)) || ranks % G) return -5;
    if (!(B == 1 || B == 8 || B == 16 || B == 32 || B == 64)) return -5;
    if (!(design->join_alg == 0 || design->join_alg == 2 || design->join_alg == 3)) return -5;
    int threads = num_threads(design->threads);
    c.R = R; c.S = S; c.nR = nR; c.nS = nS;
    c.D = (uint32_t)ranks * NUM_DPU_RANK;
    c.G = (uint32_t)G;
    c.Rg = (uint32_t)(ranks / G);
    c.B = (uint32_t)B;
    c.alg = design->join_alg;
    c.nRpart = c.Rg * NUM_DPU_RANK / c.B;
    c.tpdR = tuples_per_dpu(nR, c.D);
    c.tpdS = tuples_per_dpu(nS, c.D);
    atomic_init(&c.err, 0);

    /* which R partition each DPU holds (replication = many DPUs, one partition) */
    c.dpu_rpart = (uint32_t *)malloc(sizeof(uint32_t) * c.D);
    for (uint32_t d = 0; d < c.D; d++) c.dpu_rpart[d] = 0xFFFFFFFFu;
    uint32_t group_dpus = c.Rg * NUM_DPU_RANK;
    for (uint32_t g = 0; g < c.G; g++) {
        for (uint32_t q = 0; q < c.nRpart; q++) {
            if (c.B == 1) {
                c.dpu_rpart[g * group_dpus + q] = q;   /* RNS rankgroup_aware copy: queryHostFunc.cc:654-776 */
            } else {
                /* chipwise broadcast to the B DPUs of bank set q: queryHostFunc.cc:1181-1419 */
                uint32_t bgpr = NUM_DPU_RANK / c.B, sv = (uint32_t)ilog2u(c.B) - 3u;
                for (uint32_t b = 0; b < c.B; b++) {
                    uint32_t dd = ((q & (bgpr - 1u)) << sv) + ((q >> ilog2u(bgpr)) << 6) + lut_bg(b, c.B);
                    c.dpu_rpart[g * group_dpus + dd] = q;
                }
            }
        }
    }
    for (uint32_t d = 0; d < c.D; d++) if (c.dpu_rpart[d] == 0xFFFFFFFFu) { free(c.dpu_rpart); return -5; }

    c.cntR = (uint32_t *)calloc((size_t)c.D * c.nRpart, sizeof(uint32_t));
    c.cntS = (uint32_t *)calloc((size_t)c.D * c.D, sizeof(uint32_t));
    c.offR = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)c.D * c.nRpart);
    c.offS = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)c.D * c.D);
    c.Rp = (or_tuple_t *)malloc(sizeof(or_tuple_t) * (nR ? nR : 1));
    c.Sd = (or_tuple_t *)malloc(sizeof(or_tuple_t) * (nS ? nS : 1));
    c.rstart = (uint64_t *)calloc((size_t)c.nRpart + 1, sizeof(uint64_t));
    c.sstart = (uint64_t *)calloc((size_t)c.D + 1, sizeof(uint64_t));
    c.rescnt = (uint64_t *)calloc(c.D, sizeof(uint64_t));
    int own_res = out_cap < nS;
    c.res = own_res ? (or_pair_t *)malloc(sizeof(or_pair_t) * (nS ? nS : 1)) : out;
    int64_t ret = 0;
    if (!c.cntR || !c.cntS || !c.offR || !c.offS || !c.Rp || !c.Sd || !c.rstart || !c.sstart || !c.rescnt || !c.res) {
        ret = -2;
        goto cleanup;
    }

    /* global partition: per-source histograms (glb_partition_count.c /
     * bg_broadcast_count.c / glb_chipwise_partition_count.c), then
     * destination offsets (the page histograms of queryHostFunc.cc:1574-1769),
     * then the scatter (glb_partition_packet.c and the RNS exchange). */
    parallel_for(threads, 2ull * c.D, route_count, &c);
    {
        uint64_t run = 0;
        for (uint32_t q = 0; q < c.nRpart; q++) {
            c.rstart[q] = run;
            for (uint32_t s = 0; s < c.D; s++) { c.offR[(uint64_t)s * c.nRpart + q] = run; run += c.cntR[(uint64_t)s * c.nRpart + q]; }
        }
        c.rstart[c.nRpart] = run;
        run = 0;
        for (uint32_t d = 0; d < c.D; d++) {
            c.sstart[d] = run;
            for (uint32_t s = 0; s < c.D; s++) { c.offS[(uint64_t)s * c.D + d] = run; run += c.cntS[(uint64_t)s * c.D + d]; }
        }
        c.sstart[c.D] = run;
    }
    parallel_for(threads, 2ull * c.D, route_scatter, &c);

    /* per-DPU join */
    parallel_for(threads, c.D, dpu_join, &c);
    if (atomic_load(&c.err)) { ret = atomic_load(&c.err); goto cleanup; }

    /* result gather in DPU order, exact counts (no zero padding) */
    {
        uint64_t total = 0;
        for (uint32_t d = 0; d < c.D; d++) total += c.rescnt[d];
        if (total > out_cap) { ret = -1; goto cleanup; }
        uint64_t pos = 0;
        for (uint32_t d = 0; d < c.D; d++) {
            if (c.rescnt[d]) memmove(out + pos, c.res + c.sstart[d], sizeof(or_pair_t) * c.rescnt[d]);
            pos += c.rescnt[d];
        }
        ret = (int64_t)total;
    }
cleanup:
    free(c.dpu_rpart); free(c.cntR); free(c.cntS); free(c.offR); free(c.offS);
    free(c.Rp); free(c.Sd); free(c.rstart); free(c.sstart); free(c.rescnt);
    if (own_res) free(c.res);
    return ret;
}
