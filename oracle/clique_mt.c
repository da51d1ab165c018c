/*
 * CPU clique counter over all host cores -- TEST INFRASTRUCTURE ONLY (see
 * oracle.h): the whole-graph checker of the K3..K6 counts and the timed CPU
 * baseline of the clique benchmark. It is a CPU port of the bitmap clique
 * store's algorithm, kept deliberately plain:
 *
 *   every k-clique of a graph is a chain v0 > v1 > ... > v_{k-1} with each
 *   vertex adjacent to all earlier ones, so it is counted exactly once from
 *   its top vertex v0. For L = N+(v0) = the neighbours of v0 below v0 (d of
 *   them, sorted), row M_p holds bit q iff q < p and L[q] is a neighbour of
 *   L[p]; K3 += popc(M_p), K4 += popc(M_p & M_q) for q in M_p, K5 and K6 one
 *   and two levels deeper.
 *
 * Rows are 64-bit words built by a two-pointer merge of N+(L[p]) with
 * L[0, p). Threads take chunks of top vertices from a shared counter; each
 * keeps its own counts, summed at the end (exact uint64). The orientation is
 * by vertex id, so any labelling gives the same counts.
 */
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

typedef struct {
    int32_t n;
    const int64_t *off;
    const int32_t *adj;
    int kmax;
    int32_t first, stride;
    int64_t n_top;        /* top vertices first, first + stride, ... below n */
    atomic_llong next;    /* next index into that sequence */
    int32_t max_out;
    atomic_int failed;    /* a worker could not allocate its rows */
} job_t;

typedef struct {
    job_t *job;
    uint64_t cnt[7];
} worker_t;

static int64_t lb32(const int32_t *a, int64_t n, int32_t x) {
    int64_t lo = 0;
    while (n > 0) {
        int64_t h = n >> 1;
        if (a[lo + h] < x) {
            lo += h + 1;
            n -= h + 1;
        } else {
            n = h;
        }
    }
    return lo;
}

static void count_top(const job_t *J, int32_t v0, uint64_t *rows, uint64_t *cnt) {
    const int32_t *L = J->adj + J->off[v0];
    const int64_t d = lb32(L, J->off[v0 + 1] - J->off[v0], v0);
    const int64_t w = (d + 63) / 64;
    int64_t p, q, u;
    if (d < 2) return;
    memset(rows, 0, sizeof(uint64_t) * (size_t)(d * w));
    for (p = 1; p < d; p++) {
        const int32_t x = L[p];
        const int32_t *nx = J->adj + J->off[x];
        const int64_t xl = lb32(nx, J->off[x + 1] - J->off[x], x); /* N+(x) */
        uint64_t *mp = rows + p * w;
        int64_t i = 0, j = 0;
        while (i < xl && j < p) {
            if (nx[i] < L[j]) {
                i++;
            } else if (nx[i] > L[j]) {
                j++;
            } else {
                mp[j >> 6] |= 1ull << (j & 63);
                i++;
                j++;
            }
        }
    }
    for (p = 1; p < d; p++) {
        const uint64_t *mp = rows + p * w;
        const int64_t wp = (p >> 6) + 1;
        for (u = 0; u < wp; u++) {
            uint64_t bq = mp[u];
            cnt[3] += (uint64_t)__builtin_popcountll(bq);
            if (J->kmax < 4) continue;
            while (bq) {
                q = u * 64 + __builtin_ctzll(bq);
                bq &= bq - 1;
                const uint64_t *mq = rows + q * w;
                const int64_t wq = (q >> 6) + 1;
                int64_t a;
                for (a = 0; a < wq; a++) {
                    uint64_t br = mp[a] & mq[a];
                    cnt[4] += (uint64_t)__builtin_popcountll(br);
                    if (J->kmax < 5) continue;
                    while (br) {
                        const int64_t r = a * 64 + __builtin_ctzll(br);
                        br &= br - 1;
                        const uint64_t *mr = rows + r * w;
                        const int64_t wr = (r >> 6) + 1;
                        int64_t b;
                        for (b = 0; b < wr; b++) {
                            uint64_t bs = mp[b] & mq[b] & mr[b];
                            cnt[5] += (uint64_t)__builtin_popcountll(bs);
                            if (J->kmax < 6) continue;
                            while (bs) {
                                const int64_t s = b * 64 + __builtin_ctzll(bs);
                                bs &= bs - 1;
                                const uint64_t *ms = rows + s * w;
                                int64_t c;
                                for (c = 0; c <= (s >> 6); c++)
                                    cnt[6] += (uint64_t)__builtin_popcountll(mp[c] & mq[c] & mr[c] & ms[c]);
                            }
                        }
                    }
                }
            }
        }
    }
}

static void *worker(void *arg) {
    worker_t *W = (worker_t *)arg;
    job_t *J = W->job;
    const int64_t mo = J->max_out > 0 ? J->max_out : 1;
    uint64_t *rows = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(mo * ((mo + 63) / 64)));
    const long long chunk = 64;
    if (!rows) {
        atomic_store(&J->failed, 1);
        return NULL;
    }
    for (;;) {
        const long long i0 = atomic_fetch_add(&J->next, chunk);
        long long i;
        if (i0 >= J->n_top) break;
        for (i = i0; i < i0 + chunk && i < J->n_top; i++)
            count_top(J, (int32_t)(J->first + i * (long long)J->stride), rows, W->cnt);
    }
    free(rows);
    return NULL;
}

int orc_clique_counts_mt(int32_t n, const int64_t *off, const int32_t *adj, int kmax, int32_t first, int32_t stride,
                         int nthreads, uint64_t *counts7) {
    job_t J;
    worker_t *W;
    pthread_t *T;
    char *ok;
    int t, k;
    int32_t v;
    if (kmax < 3 || kmax > 6 || stride < 1 || first < 0 || nthreads < 1 || n < 0) return -1;
    memset(counts7, 0, sizeof(uint64_t) * 7);
    J.n = n;
    J.off = off;
    J.adj = adj;
    J.kmax = kmax;
    J.first = first;
    J.stride = stride;
    J.n_top = first < n ? ((int64_t)n - first + stride - 1) / stride : 0;
    atomic_init(&J.next, 0);
    atomic_init(&J.failed, 0);
    J.max_out = 0;
    for (v = 0; v < n; v++) {
        const int64_t d = lb32(adj + off[v], off[v + 1] - off[v], v);
        if (d > J.max_out) J.max_out = (int32_t)d;
    }
    W = (worker_t *)calloc((size_t)nthreads, sizeof(worker_t));
    T = (pthread_t *)calloc((size_t)nthreads, sizeof(pthread_t));
    ok = (char *)calloc((size_t)nthreads, 1);
    if (!W || !T || !ok) {
        free(W);
        free(T);
        free(ok);
        return -2;
    }
    /* threads that fail to start leave their share to the others */
    for (t = 0; t < nthreads; t++) {
        W[t].job = &J;
        if (t > 0) ok[t] = pthread_create(&T[t], NULL, worker, &W[t]) == 0;
    }
    worker(&W[0]);
    for (t = 1; t < nthreads; t++)
        if (ok[t]) pthread_join(T[t], NULL);
    for (t = 0; t < nthreads; t++)
        for (k = 3; k <= 6; k++) counts7[k] += W[t].cnt[k];
    free(W);
    free(T);
    free(ok);
    return atomic_load(&J.failed) ? -2 : 0;
}
