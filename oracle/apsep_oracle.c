/*
 * apsep_oracle.c -- CPU oracle for the APSEP / ANSV-left hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this file: it is the checker for tests/, for __graft_entry__.smoke()
 * and for the cpu_baseline leg of bench.py.
 *
 * Problem: for every a[i], out[i] = the largest j < i with a[j] < a[i]
 * (strict, "APSEP") or a[j] <= a[i] ("incl", PSE), else -1.
 *
 * Contents
 *   oracle_apsep            monotone stack; restates cpuApsep
 *                           (src/main.cu:102-113) and cpuPSE
 *                           (src/bench_spt_le.cu:24-35)
 *   oracle_apsep_brute      O(n^2) scan straight from the definition
 *                           (src/apsep.cuh:1-4) -- pins the stack
 *   oracle_apsep_verify     answer checker from the definition's three
 *                           clauses (sparse-table range minimum), independent
 *                           of both solvers
 *   oracle_apsep_segmented  per-segment stack over CSR offsets, answers
 *                           relative to the segment start (batched form)
 *   oracle_apsep_segmented_mt  the same on a pthread pool (CPU baseline for
 *                           the batched / served mixes)
 *   oracle_mt19937_*        std::mt19937 restated from its published
 *                           definition (Matsumoto & Nishimura 1998, the C++
 *                           standard's [rand.eng.mers] parameters); the
 *                           reference's stress generator is
 *                           src/main.cu:144-152
 *   oracle_glibc_rand_fill  glibc rand() after srand(seed) (TYPE_3 additive
 *                           feedback generator, r[i] = r[i-3] + r[i-31]),
 *                           used by src/bench_spt_le.cu:93-125
 *   oracle_glibc_intmax_fill  that generator's INT_MAX-heavy input
 *                           (src/bench_spt_le.cu:122-125)
 *   oracle_fill_*           workload generators for the benchmark classes
 *
 * Parity status: every function above is pinned by tests/test_oracle.py
 * (known-answer vectors, brute force, closed forms, libc/standard values).
 */
#define _GNU_SOURCE
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>
#include <math.h>

/* ------------------------------------------------------------------------ */
/* Solvers                                                                  */
/* ------------------------------------------------------------------------ */

/* Monotone stack.  Pop while the top does not qualify, the answer is the top
 * (or -1), push i.  Follows src/main.cu:102-113 (strict) and
 * src/bench_spt_le.cu:24-35 (incl).  `stack` must hold n entries; pass NULL
 * to allocate internally.  Returns 0 on success. */
static void stack_solve(const int32_t* a, int64_t n, int incl, int32_t* out,
                        int32_t* stk)
{
    int64_t top = 0; /* number of entries on the stack */
    if (!incl) {
        for (int64_t i = 0; i < n; i++) {
            const int32_t v = a[i];
            while (top > 0 && a[stk[top - 1]] >= v) top--;
            out[i] = top ? stk[top - 1] : -1;
            stk[top++] = (int32_t)i;
        }
    } else {
        for (int64_t i = 0; i < n; i++) {
            const int32_t v = a[i];
            while (top > 0 && a[stk[top - 1]] > v) top--;
            out[i] = top ? stk[top - 1] : -1;
            stk[top++] = (int32_t)i;
        }
    }
}

int oracle_apsep(const int32_t* a, int64_t n, int incl, int32_t* out)
{
    if (n <= 0) return 0;
    int32_t* stk = (int32_t*)malloc((size_t)n * sizeof(int32_t));
    if (!stk) return 1;
    stack_solve(a, n, incl, out, stk);
    free(stk);
    return 0;
}

/* Definition-level brute force, O(n^2): scan j = i-1 .. 0 for the first
 * qualifying element. */
int oracle_apsep_brute(const int32_t* a, int64_t n, int incl, int32_t* out)
{
    for (int64_t i = 0; i < n; i++) {
        int32_t r = -1;
        for (int64_t j = i - 1; j >= 0; j--) {
            if (incl ? (a[j] <= a[i]) : (a[j] < a[i])) { r = (int32_t)j; break; }
        }
        out[i] = r;
    }
    return 0;
}

/* Checks `out` against the definition clause by clause, without solving:
 *   (1) -1 <= out[i] < i
 *   (2) out[i] >= 0  =>  a[out[i]] qualifies for a[i]
 *   (3) nothing in (out[i], i) qualifies, i.e. min(a[out[i]+1 .. i-1]) does
 *       not qualify (sparse-table range minimum, O(n log n) memory)
 * Returns -1 if all hold, else the first failing index; -2 on OOM. */
int64_t oracle_apsep_verify(const int32_t* a, int64_t n, int incl,
                            const int32_t* out)
{
    if (n <= 0) return -1;
    int levels = 1;
    while (((int64_t)1 << levels) <= n) levels++;
    int32_t** sp = (int32_t**)calloc((size_t)levels, sizeof(int32_t*));
    if (!sp) return -2;
    int64_t fail = -1;
    for (int l = 0; l < levels; l++) {
        sp[l] = (int32_t*)malloc((size_t)n * sizeof(int32_t));
        if (!sp[l]) { fail = -2; goto done; }
    }
    memcpy(sp[0], a, (size_t)n * sizeof(int32_t));
    for (int l = 1; l < levels; l++) {
        const int64_t h = (int64_t)1 << (l - 1);
        for (int64_t i = 0; i + ((int64_t)1 << l) <= n; i++) {
            int32_t x = sp[l - 1][i], y = sp[l - 1][i + h];
            sp[l][i] = x < y ? x : y;
        }
    }
    for (int64_t i = 0; i < n; i++) {
        const int64_t r = out[i];
        if (r < -1 || r >= i) { fail = i; break; }
        if (r >= 0 && !(incl ? (a[r] <= a[i]) : (a[r] < a[i]))) { fail = i; break; }
        const int64_t lo = r + 1, hi = i - 1;
        if (lo <= hi) {
            const int64_t len = hi - lo + 1;
            int l = 0;
            while (((int64_t)1 << (l + 1)) <= len) l++;
            int32_t x = sp[l][lo], y = sp[l][hi - ((int64_t)1 << l) + 1];
            int32_t m = x < y ? x : y;
            if (incl ? (m <= a[i]) : (m < a[i])) { fail = i; break; }
        }
    }
done:
    for (int l = 0; l < levels; l++) free(sp[l]);
    free(sp);
    return fail;
}

/* Segmented form: segment s covers a[off[s] .. off[s+1]); answers are
 * relative to off[s] and never cross a segment. */
int oracle_apsep_segmented(const int32_t* a, const int64_t* off, int32_t nseg,
                           int incl, int32_t* out)
{
    int64_t maxlen = 0;
    for (int32_t s = 0; s < nseg; s++) {
        int64_t len = off[s + 1] - off[s];
        if (len < 0) return 2;
        if (len > maxlen) maxlen = len;
    }
    if (maxlen == 0) return 0;
    int32_t* stk = (int32_t*)malloc((size_t)maxlen * sizeof(int32_t));
    if (!stk) return 1;
    for (int32_t s = 0; s < nseg; s++)
        stack_solve(a + off[s], off[s + 1] - off[s], incl, out + off[s], stk);
    free(stk);
    return 0;
}

typedef struct {
    const int32_t* a; const int64_t* off; int32_t nseg; int incl; int32_t* out;
    int32_t first, stride; int rc;
} seg_job;

static void* seg_worker(void* p)
{
    seg_job* j = (seg_job*)p;
    int64_t maxlen = 0;
    for (int32_t s = j->first; s < j->nseg; s += j->stride) {
        int64_t len = j->off[s + 1] - j->off[s];
        if (len > maxlen) maxlen = len;
    }
    if (maxlen == 0) return NULL;
    int32_t* stk = (int32_t*)malloc((size_t)maxlen * sizeof(int32_t));
    if (!stk) { j->rc = 1; return NULL; }
    for (int32_t s = j->first; s < j->nseg; s += j->stride)
        stack_solve(j->a + j->off[s], j->off[s + 1] - j->off[s], j->incl,
                    j->out + j->off[s], stk);
    free(stk);
    return NULL;
}

/* Thread-pool version: segments are independent jobs, dealt round-robin. */
int oracle_apsep_segmented_mt(const int32_t* a, const int64_t* off,
                              int32_t nseg, int incl, int32_t* out,
                              int32_t nthreads)
{
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    seg_job jobs[256];
    for (int32_t t = 0; t < nthreads; t++) {
        jobs[t] = (seg_job){a, off, nseg, incl, out, t, nthreads, 0};
        if (pthread_create(&th[t], NULL, seg_worker, &jobs[t]) != 0) {
            jobs[t].rc = 3;
            seg_worker(&jobs[t]);
            th[t] = 0;
        }
    }
    int rc = 0;
    for (int32_t t = 0; t < nthreads; t++) {
        if (th[t]) pthread_join(th[t], NULL);
        if (jobs[t].rc == 1) rc = 1;
    }
    return rc;
}

/* ------------------------------------------------------------------------ */
/* std::mt19937 (32-bit Mersenne Twister, MT19937 parameters)               */
/* ------------------------------------------------------------------------ */

typedef struct { uint32_t mt[624]; int32_t idx; } oracle_mt19937;

void oracle_mt19937_seed(oracle_mt19937* s, uint32_t seed)
{
    s->mt[0] = seed;
    for (int i = 1; i < 624; i++)
        s->mt[i] = 1812433253u * (s->mt[i - 1] ^ (s->mt[i - 1] >> 30)) + (uint32_t)i;
    s->idx = 624;
}

static void mt_twist(oracle_mt19937* s)
{
    for (int i = 0; i < 624; i++) {
        uint32_t y = (s->mt[i] & 0x80000000u) | (s->mt[(i + 1) % 624] & 0x7fffffffu);
        uint32_t x = s->mt[(i + 397) % 624] ^ (y >> 1);
        if (y & 1u) x ^= 0x9908b0dfu;
        s->mt[i] = x;
    }
    s->idx = 0;
}

uint32_t oracle_mt19937_next(oracle_mt19937* s)
{
    if (s->idx >= 624) mt_twist(s);
    uint32_t y = s->mt[s->idx++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

int32_t oracle_mt19937_state_size(void) { return (int32_t)sizeof(oracle_mt19937); }

/* One trial of the reference's stress test (src/main.cu:146-152): n = 1 +
 * rng() % maxN, then n values rng() % 1000000.  Writes up to `cap` values,
 * returns n.  The state carries over between trials, as in the reference. */
int32_t oracle_stress_trial(oracle_mt19937* s, uint32_t maxN, int32_t* vals,
                            int32_t cap)
{
    int32_t n = 1 + (int32_t)(oracle_mt19937_next(s) % maxN);
    for (int32_t i = 0; i < n; i++) {
        int32_t v = (int32_t)(oracle_mt19937_next(s) % 1000000u);
        if (i < cap) vals[i] = v;
    }
    return n;
}

/* ------------------------------------------------------------------------ */
/* glibc rand() restated (TYPE_3, degree 31, separation 3)                  */
/* ------------------------------------------------------------------------ */

/* srand(seed) followed by n calls of rand(): r[0] = seed (0 -> 1),
 * r[i] = 16807 * r[i-1] mod (2^31 - 1) for i < 31 (Schrage-free 64-bit form),
 * r[31..33] = r[0..2], then r[i] = r[i-31] + r[i-3] (mod 2^32); the k-th
 * output is r[k + 344] >> 1. */
void oracle_glibc_rand_fill(uint32_t seed, int64_t n, int32_t* out)
{
    uint32_t r[34];
    int32_t w = (int32_t)seed;
    if (w == 0) w = 1;
    r[0] = (uint32_t)w;
    for (int i = 1; i < 31; i++) {
        int64_t x = (16807LL * (int32_t)r[i - 1]) % 2147483647LL;
        if (x < 0) x += 2147483647LL;
        r[i] = (uint32_t)x;
    }
    for (int i = 31; i < 34; i++) r[i] = r[i - 31];
    /* ring buffer of the last 34 values; discard the first 310 further ones */
    uint32_t ring[34];
    memcpy(ring, r, sizeof ring);
    int64_t k = 34;
    for (; k < 344; k++)
        ring[k % 34] = ring[(k - 31) % 34] + ring[(k - 3) % 34];
    for (int64_t i = 0; i < n; i++, k++) {
        uint32_t x = ring[(k - 31) % 34] + ring[(k - 3) % 34];
        ring[k % 34] = x;
        out[i] = (int32_t)(x >> 1);
    }
}

/* The INT_MAX-heavy input of src/bench_spt_le.cu:122-125: srand(seed), then
 * per element h[i] = (rand() % 4 == 0) ? INT_MAX : rand() -- the second draw
 * is made only when the first is not a multiple of 4, so element i consumes
 * one or two outputs of the same rand() stream.  Returns 0, 1 on OOM. */
int oracle_glibc_intmax_fill(uint32_t seed, int64_t n, int32_t* out)
{
    if (n <= 0) return 0;
    int32_t* r = (int32_t*)malloc((size_t)(2 * n + 1) * sizeof(int32_t));
    if (!r) return 1;
    oracle_glibc_rand_fill(seed, 2 * n + 1, r);
    int64_t k = 0;
    for (int64_t i = 0; i < n; i++) {
        if (r[k] % 4 == 0) {
            out[i] = 2147483647;
            k += 1;
        } else {
            out[i] = r[k + 1];
            k += 2;
        }
    }
    free(r);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Workload generators                                                      */
/* ------------------------------------------------------------------------ */

/* splitmix64: fast, seedable, stateless-per-index stream for the large
 * synthetic classes (uniform keys in [0, 2^31)). */
static inline uint64_t splitmix64(uint64_t x)
{
    x += 0x9e3779b97f4a7c15ULL;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31);
}

/* kind: 0 uniform [0,2^31), 1 strictly descending n-i, 2 strictly ascending
 * i, 3 uniform [0,16), 4 sawtooth i mod 4096, 5 block-descending (runs of
 * 2^16, a[i] = off[r] + (2^16-1 - i mod 2^16), off uniform in
 * [0, 2^31 - 2^16)), 6 all-equal 42, 7 uniform [0,2), 8 25% INT_MAX + 75%
 * uniform [0,2^31).  Index i is global (base + i) so chunks of a big array
 * can be filled in pieces. */
void oracle_fill(int32_t kind, uint64_t seed, int64_t base, int64_t n,
                 int64_t total, int32_t* out)
{
    for (int64_t k = 0; k < n; k++) {
        const int64_t i = base + k;
        const uint64_t h = splitmix64(seed * 0x100000001b3ULL + (uint64_t)i);
        int32_t v;
        switch (kind) {
        case 0: v = (int32_t)(h >> 33); break;
        case 1: v = (int32_t)(total - i); break;
        case 2: v = (int32_t)i; break;
        case 3: v = (int32_t)(h >> 60); break;
        case 4: v = (int32_t)(i & 4095); break;
        case 5: {
            const uint64_t r = (uint64_t)(i >> 16);
            const uint64_t ho = splitmix64(seed * 0x9e3779b1ULL + 0x5bd1e995ULL + r);
            const int64_t off = (int64_t)(ho % (uint64_t)(2147483648LL - 65536LL));
            v = (int32_t)(off + (65535 - (i & 65535)));
            break;
        }
        case 6: v = 42; break;
        case 7: v = (int32_t)(h >> 63); break;
        case 8: v = ((h & 3) == 0) ? 2147483647 : (int32_t)(h >> 33); break;
        default: v = 0; break;
        }
        out[k] = v;
    }
}

/* Log-uniform segment lengths in [lo, hi] (floor(exp(U(ln lo, ln hi)))),
 * written as CSR offsets off[0..nseg].  Returns the total. */
int64_t oracle_loguniform_offsets(uint64_t seed, int32_t nseg, double lo,
                                  double hi, int64_t* off)
{
    const double a = log(lo), b = log(hi);
    off[0] = 0;
    for (int32_t s = 0; s < nseg; s++) {
        const uint64_t h = splitmix64(seed * 0xd1b54a32d192ed03ULL + (uint64_t)s);
        const double u = (double)(h >> 11) * (1.0 / 9007199254740992.0);
        int64_t len = (int64_t)floor(exp(a + (b - a) * u));
        if (len < (int64_t)lo) len = (int64_t)lo;
        if (len > (int64_t)hi) len = (int64_t)hi;
        off[s + 1] = off[s] + len;
    }
    return off[nseg];
}
