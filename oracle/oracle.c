/*
 * CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle.h). Plain C restatement
 * of the reference's algorithms for the pattern-matching path. It is written
 * for clarity and fidelity, not speed: where the reference brute-forces
 * (permutation scans, complete-graph matching), so does this file.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define MAXN GAT_MAX_VERTEXES

/* ------------------------------------------------------------------ */
/* Pattern (include/core/pattern.hpp)                                  */
/* ------------------------------------------------------------------ */
typedef struct {
    int n;
    signed char e[MAXN * MAXN];
} pat_t;

static void pat_init(pat_t *p, int n) {
    p->n = n;
    memset(p->e, 0, sizeof(p->e));
}
static void add_edge(pat_t *p, int u, int v) { /* pattern.hpp:41-45 */
    p->e[u * p->n + v] = 1;
    p->e[v * p->n + u] = 1;
}
static int has_edge(const pat_t *p, int u, int v) { return p->e[u * p->n + v] == 1; }

/* pattern.hpp:22-39 */
static int pat_parse(const char *s, pat_t *p) {
    int len = (int)strlen(s), n;
    pat_init(p, 0);
    for (n = 1; n < len; n++) {
        if (n * n == len) {
            int i, j;
            if (n > MAXN) return -1;
            pat_init(p, n);
            for (i = 0; i < n; i++)
                for (j = 0; j < i; j++)
                    if (s[i * n + j] == '1' || s[j * n + i] == '1') add_edge(p, i, j);
            return 0;
        }
    }
    return -1;
}

/* pattern.hpp:74-82 */
static void pat_permute(const int *perm, const pat_t *p, pat_t *q) {
    int i, j;
    pat_init(q, p->n);
    for (i = 0; i < p->n; i++)
        for (j = 0; j < p->n; j++)
            if (has_edge(p, i, j)) add_edge(q, perm[i], perm[j]);
}
static int pat_eq(const pat_t *a, const pat_t *b) {
    return a->n == b->n && memcmp(a->e, b->e, (size_t)(a->n * a->n)) == 0;
}
/* pattern.hpp:85-101 */
static int pat_valid(const pat_t *p) {
    int i, j;
    for (i = 1; i < p->n; i++) {
        int ok = 0;
        for (j = 0; j < i; j++)
            if (has_edge(p, i, j)) {
                ok = 1;
                break;
            }
        if (!ok) return 0;
    }
    return 1;
}
static int pat_max_degree(const pat_t *p) { /* pattern.hpp:55-69 */
    int i, j, m = 0;
    for (i = 0; i < p->n; i++) {
        int d = 0;
        for (j = 0; j < p->n; j++) d += p->e[i * p->n + j] == 1;
        if (d > m) m = d;
    }
    return m;
}

/* std::next_permutation */
static int next_perm(int *a, int n) {
    int i = n - 2, j, t;
    while (i >= 0 && a[i] >= a[i + 1]) i--;
    if (i < 0) {
        for (i = 0, j = n - 1; i < j; i++, j--) t = a[i], a[i] = a[j], a[j] = t;
        return 0;
    }
    j = n - 1;
    while (a[j] <= a[i]) j--;
    t = a[i], a[i] = a[j], a[j] = t;
    for (i = i + 1, j = n - 1; i < j; i++, j--) t = a[i], a[i] = a[j], a[j] = t;
    return 1;
}

/* ------------------------------------------------------------------ */
/* Schedule helpers (include/core/schedule.hpp)                        */
/* ------------------------------------------------------------------ */
/* schedule.hpp:66-83 */
static int iep_suffix(const pat_t *p) {
    int k, idx;
    for (k = 2; k <= p->n - 2; k++) {
        int np = p->n - k, he = 0;
        for (idx = np + 1; idx < p->n; idx++)
            if (has_edge(p, np, idx)) {
                he = 1;
                break;
            }
        if (he) return k - 1;
    }
    return p->n - 2;
}

/* schedule.hpp:86-122 */
static int is_connected_bits(int s, int n) {
    int conn[MAXN * MAXN], cnt = 0, i, j, k;
    memset(conn, 0, sizeof(conn));
    for (i = 0; i < n; i++)
        for (j = 0; j < i; j++) {
            if (s & (1 << cnt)) conn[i * n + j] = conn[j * n + i] = 1;
            cnt++;
        }
    for (i = 0; i < n; i++) conn[i * n + i] = 1;
    for (k = 0; k < n; k++)
        for (i = 0; i < n; i++)
            for (j = 0; j < n; j++) conn[i * n + j] = conn[i * n + j] || (conn[i * n + k] && conn[k * n + j]);
    for (j = 0; j < n; j++)
        if (!conn[j]) return 0;
    return 1;
}

/* schedule.hpp:149-170 */
void orc_graph_cnt(int k, int64_t *even, int64_t *odd) {
    int n;
    int64_t e[MAXN + 1], o[MAXN + 1];
    e[1] = 1, o[1] = 0;
    for (n = 2; n <= k; n++) {
        int m = n * (n - 1) / 2, s;
        e[n] = o[n] = 0;
        for (s = 0; s < (1 << m); s++)
            if (is_connected_bits(s, n)) {
                if (__builtin_popcount((unsigned)s) % 2 == 0) e[n]++;
                else o[n]++;
            }
    }
    if (k >= 1) {
        *even = e[k];
        *odd = o[k];
    } else {
        *even = 1;
        *odd = 0;
    }
}

/* schedule.hpp:126-146 */
typedef struct {
    int a, b;
} pair_t;

static int naive_match_full(int n, const pair_t *r, int nr) {
    int data[MAXN], i, ans = 0;
    if (n >= 10) return -1;
    for (i = 0; i < n; i++) data[i] = i;
    do {
        int ok = 1;
        for (i = 0; i < nr; i++)
            if (!(data[r[i].a] < data[r[i].b])) {
                ok = 0;
                break;
            }
        ans += ok;
    } while (next_perm(data, n));
    return ans;
}

/* schedule.hpp:173-187: automorphisms in next_permutation order */
typedef struct {
    int perm[MAXN];
    int twos[MAXN][2]; /* 2-cycles in cycle order (schedule.hpp:193-213) */
    int ntwos;
    int big; /* has a cycle of length >= 3 */
} iso_t;

static void cycles_of(iso_t *t, int n) {
    int used[MAXN] = {0}, i;
    t->ntwos = 0;
    t->big = 0;
    for (i = 0; i < n; i++) {
        int len = 1, x, first2 = -1;
        if (used[i]) continue;
        used[i] = 1;
        x = t->perm[i];
        while (!used[x]) {
            used[x] = 1;
            if (len == 1) first2 = x;
            len++;
            x = t->perm[x];
        }
        if (len == 2) {
            t->twos[t->ntwos][0] = i;
            t->twos[t->ntwos][1] = first2;
            t->ntwos++;
        } else if (len >= 3) {
            /* the top-level scan stops at the first long cycle, but only the
             * fact that one exists matters to it */
            t->big = 1;
        }
    }
}

static int get_isos(const pat_t *p, iso_t **out) {
    int perm[MAXN], i, cnt = 0, cap = 16;
    iso_t *v = (iso_t *)malloc(sizeof(iso_t) * cap);
    for (i = 0; i < p->n; i++) perm[i] = i;
    do {
        pat_t q;
        pat_permute(perm, p, &q);
        if (pat_eq(&q, p)) {
            if (cnt == cap) v = (iso_t *)realloc(v, sizeof(iso_t) * (cap *= 2));
            memcpy(v[cnt].perm, perm, sizeof(perm));
            cycles_of(&v[cnt], p->n);
            cnt++;
        }
    } while (next_perm(perm, p->n));
    *out = v;
    return cnt;
}

int64_t orc_automorphisms(const char *adj) {
    pat_t p;
    iso_t *v;
    int c;
    if (pat_parse(adj, &p)) return -1;
    c = get_isos(&p, &v);
    free(v);
    return c;
}

/* schedule.hpp:215-237 (Kahn's algorithm, duplicates allowed) */
static int is_dag(const pair_t *e, int ne, int size) {
    int degree[MAXN] = {0}, q[MAXN * MAXN], qh = 0, qt = 0, visit = 0, i, k;
    for (i = 0; i < ne; i++) degree[e[i].b]++;
    for (i = 0; i < size; i++)
        if (degree[i] == 0) q[qt++] = i;
    while (qh < qt) {
        int x = q[qh++];
        visit++;
        for (k = 0; k < ne; k++)
            if (e[k].a == x)
                if (--degree[e[k].b] == 0) q[qt++] = e[k].b;
    }
    return visit == size;
}

/* growable list of restriction sets */
typedef struct {
    pair_t *pairs;
    int *start; /* set i = pairs[start[i] .. start[i+1]) */
    int nsets, npairs, cap_sets, cap_pairs;
} rlist_t;

static void rlist_push(rlist_t *L, const pair_t *p, int n) {
    int i;
    if (L->nsets + 2 > L->cap_sets) {
        L->cap_sets = L->cap_sets ? L->cap_sets * 2 : 16;
        L->start = (int *)realloc(L->start, sizeof(int) * L->cap_sets);
        if (L->nsets == 0) L->start[0] = 0;
    }
    if (L->npairs + n > L->cap_pairs) {
        while (L->npairs + n > L->cap_pairs) L->cap_pairs = L->cap_pairs ? L->cap_pairs * 2 : 64;
        L->pairs = (pair_t *)realloc(L->pairs, sizeof(pair_t) * L->cap_pairs);
    }
    for (i = 0; i < n; i++) L->pairs[L->npairs++] = p[i];
    L->nsets++;
    L->start[L->nsets] = L->npairs;
}

/* schedule.hpp:239-293. `idx` lists the surviving automorphisms (indices
 * into isos) in order. */
static void aggressive_dfs(const pair_t *base, int nb, const iso_t *isos, const int *idx, int nidx, rlist_t *out,
                           int size) {
    int *keep = (int *)malloc(sizeof(int) * (nidx + 1)), nk = 0, i, c, k;
    pair_t *test = (pair_t *)malloc(sizeof(pair_t) * (2 * nb + 1));
    for (i = 0; i < nidx; i++) {
        const int *iso = isos[idx[i]].perm;
        for (k = 0; k < nb; k++) test[k] = base[k];
        for (k = 0; k < nb; k++) {
            test[nb + k].a = iso[base[k].a];
            test[nb + k].b = iso[base[k].b];
        }
        if (is_dag(test, 2 * nb, size)) keep[nk++] = idx[i];
    }
    free(test);
    if (nk == 1) {
        rlist_push(out, base, nb);
        free(keep);
        return;
    }
    for (i = 0; i < nk; i++) {
        const iso_t *g = &isos[keep[i]];
        for (c = 0; c < g->ntwos; c++) {
            int *nidx2 = (int *)malloc(sizeof(int) * (nk + 1)), n2 = 0;
            pair_t *nb2 = (pair_t *)malloc(sizeof(pair_t) * (nb + 1));
            for (k = 0; k < nk; k++)
                if (k != i) nidx2[n2++] = keep[k];
            for (k = 0; k < nb; k++) nb2[k] = base[k];
            nb2[nb].a = g->twos[c][0];
            nb2[nb].b = g->twos[c][1];
            aggressive_dfs(nb2, nb + 1, isos, nidx2, n2, out, size);
            free(nidx2);
            free(nb2);
        }
        /* the reference's early-exit counter is never incremented here */
    }
    free(keep);
}

/* schedule.hpp:296-360 */
static void restricts_generate(const pat_t *p, rlist_t *valid) {
    iso_t *isos;
    int niso = get_isos(p, &isos), i, ans;
    int *rest = (int *)malloc(sizeof(int) * niso), nrest = 0;
    pair_t ordered[MAXN * MAXN];
    int nordered = 0;
    rlist_t all = {0};
    for (i = 0; i < niso; i++) {
        int two = isos[i].big ? -1 : isos[i].ntwos;
        if (two == 1) {
            ordered[nordered].a = isos[i].twos[0][0];
            ordered[nordered].b = isos[i].twos[0][1];
            nordered++;
        } else {
            rest[nrest++] = i;
        }
    }
    aggressive_dfs(ordered, nordered, isos, rest, nrest, &all, p->n);
    ans = naive_match_full(p->n, NULL, 0) / niso;
    for (i = 0; i < all.nsets; i++) {
        int cur = naive_match_full(p->n, all.pairs + all.start[i], all.start[i + 1] - all.start[i]);
        if (cur == ans) rlist_push(valid, all.pairs + all.start[i], all.start[i + 1] - all.start[i]);
    }
    free(all.pairs);
    free(all.start);
    free(rest);
    free(isos);
}

static int pair_cmp(const void *x, const void *y) {
    const pair_t *a = (const pair_t *)x, *b = (const pair_t *)y;
    if (a->a != b->a) return a->a - b->a;
    return a->b - b->b;
}

/* schedule.hpp:362-437, evaluation order kept */
static double estimate_restrictions(const pat_t *p, const pair_t *restrictions, int nr, int v_cnt, int64_t e_cnt,
                                    int64_t tri_cnt) {
    int size = p->n, iep_suffix_num = iep_suffix(p), i, j;
    pair_t r[MAXN * MAXN];
    double p0 = e_cnt * 1.0 / v_cnt / v_cnt;
    double p1 = tri_cnt * 1.0 * v_cnt / e_cnt / e_cnt;
    double sum[MAXN * MAXN], total = 1, val = 1;
    int tmp[MAXN];
    int inv[MAXN][MAXN], ninv[MAXN];
    memcpy(r, restrictions, sizeof(pair_t) * nr);
    qsort(r, (size_t)nr, sizeof(pair_t), pair_cmp);
    for (i = 0; i < nr; i++) sum[i] = 0;
    for (i = 0; i < size; i++) tmp[i] = i;
    do {
        for (i = 0; i < nr; i++)
            if (tmp[r[i].a] > tmp[r[i].b]) sum[i] += 1;
            else break;
    } while (next_perm(tmp, size));
    for (i = 2; i <= size; i++) total *= i;
    for (i = 0; i < nr; i++) sum[i] = sum[i] / total;
    for (i = nr - 1; i > 0; --i) sum[i] /= sum[i - 1];
    for (i = 0; i < size; i++) ninv[i] = 0;
    for (i = size - iep_suffix_num - 1; i >= 0; --i) {
        int cnt_forward = 0, c;
        for (j = 0; j < i; j++)
            if (has_edge(p, j, i)) ++cnt_forward;
        c = cnt_forward;
        for (j = i - 1; j >= 0; --j)
            if (has_edge(p, j, i)) inv[j][ninv[j]++] = c--;
        for (j = 0; j < ninv[i]; j++)
            if (inv[i][j] > 1)
                val += v_cnt * p0 * pow(p1, (double)(inv[i][j] - 2)) * log2(p0 * v_cnt);
        for (j = 0; j < nr; j++)
            if (r[j].b == i) val *= sum[j];
        if (i)
            val *= v_cnt * p0 * pow(p1, (double)(cnt_forward - 1));
        else
            val *= v_cnt;
    }
    return val;
}

/* IEP groups (schedule.hpp:526-590) */
typedef struct {
    int ngroups;
    int nblocks[64];
    int blocks[64][MAXN][MAXN]; /* suffix-vertex ids */
    int blen[64][MAXN];
    int coef[64];
} iep_t;

static void iep_rec(int depth, int *id, int gc, int suffix, iep_t *out, const int64_t *ge, const int64_t *go) {
    int i;
    if (depth == suffix) {
        int gsize[MAXN] = {0}, g, v;
        int64_t f = 1, s = 0;
        int k = out->ngroups++;
        for (i = 0; i < suffix; i++) gsize[id[i]]++;
        for (i = 0; i < gc; i++) {
            int64_t nf = f * ge[gsize[i]] + s * go[gsize[i]];
            int64_t ns = f * go[gsize[i]] + s * ge[gsize[i]];
            f = nf, s = ns;
        }
        out->nblocks[k] = gc;
        for (g = 0; g < gc; g++) {
            out->blen[k][g] = 0;
            for (v = 0; v < suffix; v++)
                if (id[v] == g) out->blocks[k][g][out->blen[k][g]++] = v;
        }
        out->coef[k] = (int)(f - s);
        return;
    }
    id[depth] = gc;
    iep_rec(depth + 1, id, gc + 1, suffix, out, ge, go);
    for (i = 0; i < gc; i++) {
        id[depth] = i;
        iep_rec(depth + 1, id, gc, suffix, out, ge, go);
    }
}

/* ------------------------------------------------------------------ */
/* Schedule (schedule.hpp:592-1026)                                    */
/* ------------------------------------------------------------------ */
typedef struct {
    int np;
    int data[GAT_MAX_PREFIXS + 8][MAXN];
    int depth[GAT_MAX_PREFIXS + 8];
    int father[GAT_MAX_PREFIXS + 8];
} trie_t;

static int find_prefix(const trie_t *t, const int *d, int n) {
    int i;
    for (i = 0; i < t->np; i++)
        if (t->depth[i] == n && memcmp(t->data[i], d, sizeof(int) * n) == 0) return i;
    return t->np;
}

/* schedule.hpp:655-675 */
static int insert_prefix(trie_t *t, const int *d, int n) {
    int pos, f;
    if (n == 0) return -1;
    pos = find_prefix(t, d, n);
    if (pos != t->np) return pos;
    f = insert_prefix(t, d, n - 1);
    if (t->np >= GAT_MAX_PREFIXS + 8) return -2;
    memcpy(t->data[t->np], d, sizeof(int) * n);
    t->depth[t->np] = n;
    t->father[t->np] = f;
    return t->np++;
}

int orc_schedule(const char *adj, gat_schedule_data *o) {
    pat_t p, best_p;
    int perm[MAXN], best_perm[MAXN], i, j, k;
    int best_iep = -1, have_best = 0;
    double best_cost = 1e18;
    pair_t best_r[MAXN * MAXN];
    int nbest_r = 0;
    int niso_total;
    iso_t *isos_in;

    if (pat_parse(adj, &p) || p.n < 3) return -1;
    memset(o, 0, sizeof(*o));
    niso_total = get_isos(&p, &isos_in);
    free(isos_in);

    for (i = 0; i < p.n; i++) perm[i] = i;
    do { /* schedule.hpp:878-926 */
        pat_t q;
        int suf;
        rlist_t rl = {0};
        pat_permute(perm, &p, &q);
        if (!pat_valid(&q)) continue;
        suf = iep_suffix(&q);
        if (suf < best_iep) continue;
        if (suf > best_iep) {
            best_iep = suf;
            best_cost = 1e18;
            have_best = 0;
        }
        restricts_generate(&q, &rl);
        for (k = 0; k < rl.nsets; k++) {
            int nr = rl.start[k + 1] - rl.start[k];
            double est = estimate_restrictions(&q, rl.pairs + rl.start[k], nr, 5000, 1000000, 10000);
            if (!have_best || est < best_cost) {
                have_best = 1;
                best_cost = est;
                memcpy(best_perm, perm, sizeof(perm));
                memcpy(best_r, rl.pairs + rl.start[k], sizeof(pair_t) * nr);
                nbest_r = nr;
            }
        }
        free(rl.pairs);
        free(rl.start);
    } while (next_perm(perm, p.n));
    if (!have_best) return -1;
    pat_permute(best_perm, &p, &best_p);

    {
        /* calculate_pattern (schedule.hpp:677-750) */
        const pat_t *q = &best_p;
        int suffix = iep_suffix(q), basic = q->n - suffix;
        int64_t ge[MAXN + 1], go[MAXN + 1];
        int id[MAXN];
        iep_t *iep = (iep_t *)calloc(1, sizeof(iep_t));
        trie_t t;
        int vid_prefix[GAT_MAX_PREFIXS + 8], nvid = 0, nsub = 0;
        int ans_pos[GAT_MAX_IEP_GROUPS * 2], coefs[GAT_MAX_IEP_GROUPS * 2], flags[GAT_MAX_IEP_GROUPS * 2];
        int order[GAT_MAX_PREFIXS + 8], where[GAT_MAX_PREFIXS + 8];
        pair_t fr[MAXN * MAXN];
        int nfr = 0, cs, np;

        t.np = 0;
        ge[0] = 1, go[0] = 0;
        for (k = 1; k <= suffix; k++) orc_graph_cnt(k, &ge[k], &go[k]);
        iep_rec(0, id, 0, suffix, iep, ge, go);
        for (i = 0; i < basic; i++) {
            int d[MAXN], n = 0;
            for (j = 0; j < i; j++)
                if (has_edge(q, i, j)) d[n++] = j;
            if (insert_prefix(&t, d, n) == -2) return -2;
        }
        for (k = 0; k < iep->ngroups; k++) {
            int g;
            for (g = 0; g < iep->nblocks[k]; g++) {
                int d[MAXN], n = 0, pid, vid;
                for (i = 0; i < basic; i++)
                    for (j = 0; j < iep->blen[k][g]; j++)
                        if (has_edge(q, i, basic + iep->blocks[k][g][j])) {
                            d[n++] = i;
                            break;
                        }
                pid = insert_prefix(&t, d, n);
                if (pid == -2) return -2;
                for (vid = 0; vid < nvid; vid++)
                    if (vid_prefix[vid] == pid) break;
                if (vid == nvid) vid_prefix[nvid++] = pid;
                if (nsub >= GAT_MAX_IEP_GROUPS) return -2;
                ans_pos[nsub] = vid;
                if (g == iep->nblocks[k] - 1) {
                    coefs[nsub] = iep->coef[k];
                    flags[nsub] = 1;
                } else {
                    coefs[nsub] = 0;
                    flags[nsub] = 0;
                }
                nsub++;
            }
        }
        free(iep);
        np = t.np;
        if (np > GAT_MAX_PREFIXS) return -2;
        /* sort_prefixs (schedule.hpp:614-652): stable by last element */
        for (i = 0; i < np; i++) order[i] = i;
        for (i = 1; i < np; i++) { /* insertion sort by (last, index) */
            int x = order[i], y = i - 1;
            while (y >= 0 && (t.data[order[y]][t.depth[order[y]] - 1] > t.data[x][t.depth[x] - 1])) {
                order[y + 1] = order[y];
                y--;
            }
            order[y + 1] = x;
        }
        for (i = 0; i < np; i++) where[order[i]] = i;
        o->v_cnt = q->n;
        o->iep_suffix_vertexes = suffix;
        o->basic_vertexes = basic;
        o->total_prefix_num = np;
        for (i = 0; i < GAT_MAX_PREFIXS; i++) o->prefixs_father[i] = -1;
        for (i = 0; i < np; i++) {
            int src = order[i];
            o->prefix_depth[i] = t.depth[src];
            memcpy(o->prefix_data[i], t.data[src], sizeof(int) * t.depth[src]);
            o->prefixs_father[i] = t.father[src] < 0 ? -1 : where[t.father[src]];
        }
        o->iep_prefix_num = nvid;
        for (i = 0; i < nvid; i++) o->iep_vertex_id[i] = where[vid_prefix[i]];
        o->subgroups_num = nsub;
        for (i = 0; i < nsub; i++) {
            o->iep_ans_pos[i] = ans_pos[i];
            o->iep_coef[i] = coefs[i];
            o->iep_flag[i] = flags[i];
        }
        /* build_loop_invariant (schedule.hpp:786-818) */
        o->loop_set_prefix_id[0] = -1;
        for (i = 1; i < basic; i++) {
            int d[MAXN], n = 0, pos;
            for (j = 0; j < i; j++)
                if (has_edge(q, j, i)) d[n++] = j;
            for (pos = 0; pos < np; pos++)
                if (o->prefix_depth[pos] == n && memcmp(o->prefix_data[pos], d, sizeof(int) * n) == 0) break;
            if (pos == np) return -3;
            o->loop_set_prefix_id[i] = pos;
        }
        {
            int cur = 0;
            for (i = 0; i < basic; i++) {
                o->vertex_prefix_start[i] = cur;
                while (cur < np && o->prefix_data[cur][o->prefix_depth[cur] - 1] == i) cur++;
            }
            o->vertex_prefix_start[basic] = np;
        }
        /* build_restrictions (schedule.hpp:820-845): sort by (second, first),
         * drop second >= basic; the start scan stops at the list end */
        for (i = 0; i < nbest_r; i++)
            if (best_r[i].b < basic) fr[nfr++] = best_r[i];
        for (i = 1; i < nfr; i++) {
            pair_t x = fr[i];
            int y = i - 1;
            while (y >= 0 && (fr[y].b > x.b || (fr[y].b == x.b && fr[y].a > x.a))) {
                fr[y + 1] = fr[y];
                y--;
            }
            fr[y + 1] = x;
        }
        cs = 0;
        for (i = 0; i < basic; i++) {
            while (cs < nfr && fr[cs].b < i) cs++;
            o->restrictions_start[i] = cs;
        }
        o->restrictions_start[basic] = nfr;
        o->total_restrictions_num = nfr;
        for (i = 0; i < nfr; i++) {
            o->restrictions[i] = fr[i].a;
            o->restriction_second[i] = fr[i].b;
        }
        /* build_only_need_size (schedule.hpp:847-862) */
        {
            int has_child[GAT_MAX_PREFIXS] = {0};
            for (i = 0; i < np; i++)
                if (o->prefixs_father[i] != -1) has_child[o->prefixs_father[i]] = 1;
            for (i = 0; i < np; i++) o->prefixs_size_only[i] = has_child[i] ? 0 : 1;
        }
        for (i = 0; i < q->n; i++) {
            unsigned row = 0;
            o->perm[i] = best_perm[i];
            for (j = 0; j < q->n; j++)
                if (has_edge(q, i, j)) row |= 1u << j;
            o->adj_rows[i] = row;
        }
        o->n_all_restrictions = nbest_r;
        for (i = 0; i < nbest_r; i++) {
            o->all_restr_first[i] = best_r[i].a;
            o->all_restr_second[i] = best_r[i].b;
        }
        o->automorphisms = niso_total;
        o->best_cost = best_cost;
    }
    (void)pat_max_degree;
    return 0;
}

/* ------------------------------------------------------------------ */
/* Data-graph helpers                                                  */
/* ------------------------------------------------------------------ */
static int lower_bound_i32(const int32_t *a, int64_t n, int32_t x) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        int64_t mid = (lo + hi) / 2;
        if (a[mid] < x) lo = mid + 1;
        else hi = mid;
    }
    return (int)lo;
}
/* graph_backend.cuh:140-147 */
static int g_has_edge(const int64_t *off, const int32_t *adj, int32_t u, int32_t v) {
    int64_t du = off[u + 1] - off[u], dv = off[v + 1] - off[v];
    int64_t pos;
    if (du > dv) {
        int32_t t = u;
        u = v;
        v = t;
        du = dv;
    }
    pos = lower_bound_i32(adj + off[u], du, v);
    return pos != du && adj[off[u] + pos] == v;
}

/* naive_pm.cu:10-44 (uint64 accumulator instead of int) */
static uint64_t naive_rec(int depth, int32_t *chosen, int32_t n, const int64_t *off, const int32_t *adj,
                          const pat_t *p) {
    uint64_t ans = 0;
    int32_t i;
    int j;
    if (depth == p->n) return 1;
    for (i = 0; i < n; i++) {
        int flag = 1;
        for (j = 0; j < depth; j++) {
            int in_p = has_edge(p, j, depth);
            int in_g = g_has_edge(off, adj, chosen[j], i);
            if (chosen[j] == i || (in_p && !in_g)) {
                flag = 0;
                break;
            }
        }
        if (flag) {
            chosen[depth] = i;
            ans += naive_rec(depth + 1, chosen, n, off, adj, p);
        }
    }
    return ans;
}

uint64_t orc_naive_count(int32_t n, const int64_t *off, const int32_t *adj, const char *pattern) {
    pat_t p;
    int32_t chosen[MAXN];
    if (pat_parse(pattern, &p)) return 0;
    return naive_rec(0, chosen, n, off, adj, &p);
}

/* Labeled embeddings by backtracking over a BFS order of the pattern. */
typedef struct {
    const int64_t *off;
    const int32_t *adj;
    const pat_t *p;
    int order[MAXN], anchor[MAXN];
    int32_t map[MAXN];
} lab_t;

static uint64_t lab_rec(lab_t *L, int i) {
    uint64_t ans = 0;
    int u, j;
    int64_t e;
    int32_t a;
    if (i == L->p->n) return 1;
    u = L->order[i];
    a = L->map[L->anchor[i]];
    for (e = L->off[a]; e < L->off[a + 1]; e++) {
        int32_t w = L->adj[e];
        int ok = 1;
        for (j = 0; j < i && ok; j++) {
            if (L->map[L->order[j]] == w) ok = 0;
            else if (has_edge(L->p, u, L->order[j]) && !g_has_edge(L->off, L->adj, w, L->map[L->order[j]])) ok = 0;
        }
        if (ok) {
            L->map[u] = w;
            ans += lab_rec(L, i + 1);
        }
    }
    return ans;
}

uint64_t orc_labeled_count(int32_t n, const int64_t *off, const int32_t *adj, const char *pattern) {
    pat_t p;
    lab_t L;
    int seen[MAXN] = {0}, qh = 0, qt = 0, i, j;
    uint64_t ans = 0;
    int32_t v;
    if (pat_parse(pattern, &p) || p.n < 1) return 0;
    L.off = off;
    L.adj = adj;
    L.p = &p;
    L.order[qt++] = 0;
    seen[0] = 1;
    L.anchor[0] = -1;
    while (qh < qt) {
        int x = L.order[qh++];
        for (j = 0; j < p.n; j++)
            if (!seen[j] && has_edge(&p, x, j)) {
                seen[j] = 1;
                L.anchor[qt] = x;
                L.order[qt++] = j;
            }
    }
    if (qt != p.n) return 0; /* disconnected pattern */
    for (i = 0; i < p.n; i++) L.map[i] = -1;
    for (v = 0; v < n; v++) {
        L.map[0] = v;
        ans += lab_rec(&L, 1);
    }
    return ans;
}

/* ------------------------------------------------------------------ */
/* Restricted maps of a Schedule-ordered pattern, by root              */
/* ------------------------------------------------------------------ */
/* Maps psi of the pattern (vertex i adjacent to j iff bit j of rows[i])
 * into the graph that are injective, edge-preserving and satisfy
 * psi(second) < psi(first) for every restriction, the direction the
 * reference engine tests (include/engine/engine_kernel.cuh:321-334). The
 * count does not depend on the order the vertices are placed in, so after
 * vertex 0 (the root) the next vertex is always one with the most placed
 * neighbours (ties: lowest index), which keeps a low-degree root from
 * walking a hub's list twice. Level i takes the list of its placed
 * neighbour with the fewest entries and cuts it to the open interval the
 * restrictions between i and placed levels leave; each restriction is
 * thereby tested once, at its later level. */
typedef struct {
    const int64_t *off;
    const int32_t *adj;
    int32_t n;
    int k;
    uint32_t rows[MAXN];
    int nr;
    int rf[MAXN * MAXN], rs[MAXN * MAXN];
    int32_t map[MAXN];
} rst_t;

static uint64_t rst_rec(rst_t *R, int i) {
    int j, r, anchor = -1;
    int32_t lo = -1, hi = R->n; /* candidates w with lo < w < hi */
    int64_t e, e0, e1;
    uint64_t ans = 0;
    for (j = 0; j < i; j++)
        if ((R->rows[i] >> j) & 1u) {
            int32_t x = R->map[j], y = anchor >= 0 ? R->map[anchor] : 0;
            if (anchor < 0 || R->off[x + 1] - R->off[x] < R->off[y + 1] - R->off[y]) anchor = j;
        }
    for (r = 0; r < R->nr; r++) {
        if (R->rs[r] == i && R->rf[r] < i && R->map[R->rf[r]] < hi) hi = R->map[R->rf[r]];
        if (R->rf[r] == i && R->rs[r] < i && R->map[R->rs[r]] > lo) lo = R->map[R->rs[r]];
    }
    if (anchor >= 0) {
        e0 = R->off[R->map[anchor]];
        e1 = R->off[R->map[anchor] + 1];
    } else { /* no earlier neighbour: every vertex is a candidate */
        e0 = 0;
        e1 = R->n;
    }
    for (e = e0; e < e1; e++) {
        int32_t w = anchor >= 0 ? R->adj[e] : (int32_t)e;
        int ok = 1;
        if (w <= lo) continue;
        if (w >= hi) break;
        for (j = 0; j < i && ok; j++) {
            if (R->map[j] == w) ok = 0;
            else if (j != anchor && ((R->rows[i] >> j) & 1u) && !g_has_edge(R->off, R->adj, w, R->map[j])) ok = 0;
        }
        if (!ok) continue;
        if (i == R->k - 1) {
            ans++;
        } else {
            R->map[i] = w;
            ans += rst_rec(R, i + 1);
        }
    }
    return ans;
}

uint64_t orc_restricted_count(int32_t n, const int64_t *off, const int32_t *adj, int k, const uint32_t *rows,
                              int nr, const int32_t *first, const int32_t *second, int32_t root_lo,
                              int32_t root_hi) {
    rst_t R;
    uint64_t ans = 0;
    int32_t v;
    int i;
    int order[MAXN], pos[MAXN], placed = 1, j;
    if (k < 1 || k > MAXN || nr < 0 || nr > MAXN * MAXN) return 0;
    R.off = off;
    R.adj = adj;
    R.n = n;
    R.k = k;
    R.nr = nr;
    for (i = 0; i < k; i++) pos[i] = -1;
    order[0] = 0;
    pos[0] = 0;
    while (placed < k) {
        int best = -1, best_links = -1;
        for (i = 0; i < k; i++) {
            int links = 0;
            if (pos[i] >= 0) continue;
            for (j = 0; j < k; j++)
                if (pos[j] >= 0 && (((rows[i] >> j) | (rows[j] >> i)) & 1u)) links++;
            if (links > best_links) {
                best = i;
                best_links = links;
            }
        }
        pos[best] = placed;
        order[placed++] = best;
    }
    for (i = 0; i < k; i++) {
        R.rows[i] = 0;
        for (j = 0; j < k; j++)
            if (((rows[order[i]] >> order[j]) | (rows[order[j]] >> order[i])) & 1u) R.rows[i] |= 1u << j;
    }
    for (i = 0; i < nr; i++) {
        R.rf[i] = pos[first[i]];
        R.rs[i] = pos[second[i]];
    }
    if (root_lo < 0) root_lo = 0;
    if (root_hi > n) root_hi = n;
    for (v = root_lo; v < root_hi; v++) {
        if (k == 1) {
            ans++;
            continue;
        }
        R.map[0] = v;
        ans += rst_rec(&R, 1);
    }
    return ans;
}

/* ------------------------------------------------------------------ */
/* Reference engine semantics                                          */
/* ------------------------------------------------------------------ */
typedef struct {
    const int64_t *off;
    const int32_t *adj;
    int32_t n;
    const gat_schedule_data *s;
    int32_t *buf[GAT_MAX_PREFIXS];
    const int32_t *ptr[GAT_MAX_PREFIXS];
    int64_t len[GAT_MAX_PREFIXS];
    int32_t chosen[MAXN];
    uint64_t total;
} ref_t;

static int in_sorted(const int32_t *a, int64_t n, int32_t x) {
    int64_t pos = lower_bound_i32(a, n, x);
    return pos < n && a[pos] == x;
}

/* get_iep_answer_kernel (engine_kernel.cuh:343-432) */
static void ref_iep(ref_t *R) {
    const gat_schedule_data *s = R->s;
    int64_t ans[GAT_MAX_PREFIXS];
    int k, i, gid, last = -1;
    uint64_t val = 1, local = 0;
    for (k = 0; k < s->iep_prefix_num; k++) {
        int pid = s->iep_vertex_id[k];
        if (s->prefixs_size_only[pid]) {
            ans[k] = R->len[pid];
        } else { /* subtraction_size_onethread (array_vertex_set.cuh:318-331) */
            int64_t sub = 0;
            for (i = 0; i < s->basic_vertexes; i++) sub += in_sorted(R->ptr[pid], R->len[pid], R->chosen[i]);
            ans[k] = R->len[pid] - sub;
        }
    }
    for (gid = 0; gid < s->subgroups_num; gid++) {
        uint64_t a = (uint64_t)ans[s->iep_ans_pos[gid]];
        if (gid == last + 1) val = a;
        else val *= a;
        if (s->iep_flag[gid]) {
            last = gid;
            local += val * (uint64_t)(int64_t)s->iep_coef[gid];
        }
    }
    R->total += local;
}

/* prefix sets that end at vertex l, for chosen vertex v
 * (extend_p_storage_kernel, engine_kernel.cuh:173-286) */
static void ref_build(ref_t *R, int l, int32_t v) {
    const gat_schedule_data *s = R->s;
    int p, i, ap = 0;
    const int32_t *nv = R->adj + R->off[v];
    int64_t dv = R->off[v + 1] - R->off[v];
    for (i = 0; i < l; i++) ap |= R->chosen[i] == v; /* has_data<l> */
    for (p = s->vertex_prefix_start[l]; p < s->vertex_prefix_start[l + 1]; p++) {
        int f = s->prefixs_father[p];
        if (ap) {
            R->len[p] = 0;
            continue;
        }
        if (f == -1) {
            R->ptr[p] = nv;
            R->len[p] = dv;
        } else if (!s->prefixs_size_only[p]) {
            int64_t k, m = 0;
            for (k = 0; k < dv; k++)
                if (in_sorted(R->ptr[f], R->len[f], nv[k])) R->buf[p][m++] = nv[k];
            R->ptr[p] = R->buf[p];
            R->len[p] = m;
        } else { /* intersect_size<l>: excludes chosen[0..l-1] */
            int64_t k, m = 0;
            for (k = 0; k < dv; k++) {
                int ex = 0;
                for (i = 0; i < l; i++) ex |= R->chosen[i] == nv[k];
                if (!ex && in_sorted(R->ptr[f], R->len[f], nv[k])) m++;
            }
            R->len[p] = m;
        }
    }
}

static void ref_level(ref_t *R, int l) {
    const gat_schedule_data *s = R->s;
    int lsp, i;
    int32_t minv, k, cnt;
    const int32_t *loop;
    if (l == s->basic_vertexes - 1) {
        ref_iep(R);
        return;
    }
    /* prepare_v_storage_kernel (engine_kernel.cuh:290-338) */
    lsp = s->loop_set_prefix_id[l + 1];
    minv = R->n;
    for (i = s->restrictions_start[l + 1]; i < s->restrictions_start[l + 2]; i++)
        if (R->chosen[s->restrictions[i]] < minv) minv = R->chosen[s->restrictions[i]];
    loop = R->ptr[lsp];
    cnt = lower_bound_i32(loop, R->len[lsp], minv);
    for (k = 0; k < cnt; k++) {
        int32_t v = loop[k];
        R->chosen[l + 1] = v;
        ref_build(R, l + 1, v);
        ref_level(R, l + 1);
    }
}

uint64_t orc_reference_count(int32_t n, const int64_t *off, const int32_t *adj, const gat_schedule_data *s,
                             int *ok) {
    return orc_reference_count_roots(n, off, adj, s, 0, n, ok);
}

uint64_t orc_reference_count_roots(int32_t n, const int64_t *off, const int32_t *adj, const gat_schedule_data *s,
                                   int32_t root_lo, int32_t root_hi, int *ok) {
    ref_t R;
    int p, l;
    int32_t maxd = 0, v;
    memset(&R, 0, sizeof(R));
    *ok = 1;
    for (l = 1; l < s->basic_vertexes; l++) {
        int pid = s->loop_set_prefix_id[l];
        if (s->prefixs_size_only[pid] && s->prefix_depth[pid] > 1) *ok = 0;
    }
    if (!*ok) return 0;
    for (v = 0; v < n; v++)
        if (off[v + 1] - off[v] > maxd) maxd = (int32_t)(off[v + 1] - off[v]);
    for (p = 0; p < s->total_prefix_num; p++) R.buf[p] = (int32_t *)malloc(sizeof(int32_t) * (maxd + 1));
    R.off = off;
    R.adj = adj;
    R.n = n;
    R.s = s;
    /* perform_search + first_extend_kernel (engine.cuh:232-265,
     * engine_kernel.cuh:40-60): every first vertex starts its own walk */
    if (root_lo < 0) root_lo = 0;
    if (root_hi > n) root_hi = n;
    for (v = root_lo; v < root_hi; v++) {
        R.chosen[0] = v;
        ref_build(&R, 0, v);
        ref_level(&R, 0);
    }
    for (p = 0; p < s->total_prefix_num; p++) free(R.buf[p]);
    return R.total;
}

/* ------------------------------------------------------------------ */
/* Oriented clique walk with algorithmic-byte accounting               */
/* ------------------------------------------------------------------ */
typedef struct {
    const int64_t *off;
    const int32_t *adj;
    int k;
    int32_t *S[MAXN];
    int64_t Slen[MAXN];
    uint64_t count, bytes;
} clq_t;

/* S holds the candidates of depth d+1 (all below v_d); v = v_d */
static void clq_rec(clq_t *Q, int d) {
    int64_t i;
    for (i = 0; i < Q->Slen[d - 1]; i++) {
        int32_t v = Q->S[d - 1][i];
        const int32_t *nv = Q->adj + Q->off[v];
        int64_t a = lower_bound_i32(nv, Q->off[v + 1] - Q->off[v], v);
        int64_t b = lower_bound_i32(Q->S[d - 1], Q->Slen[d - 1], v), j, m = 0;
        for (j = 0; j < a; j++)
            if (in_sorted(Q->S[d - 1], b, nv[j])) {
                if (d < Q->k - 2) Q->S[d][m] = nv[j];
                m++;
            }
        if (d == Q->k - 2) {
            Q->count += (uint64_t)m;
            Q->bytes += 4ull * (uint64_t)(a + b) + 16ull;
        } else {
            Q->bytes += 4ull * (uint64_t)(a + b) + 4ull * (uint64_t)m + 16ull;
            Q->Slen[d] = m;
            clq_rec(Q, d + 1);
        }
    }
}

uint64_t orc_clique_walg(int32_t n, const int64_t *off, const int32_t *adj, int k, uint64_t *walg) {
    return orc_clique_walg_sample(n, off, adj, k, 1, 0, walg);
}

uint64_t orc_clique_walg_sample(int32_t n, const int64_t *off, const int32_t *adj, int k, int32_t stride,
                                int32_t first, uint64_t *walg) {
    clq_t Q;
    int32_t v0, maxd = 0, v;
    int d;
    memset(&Q, 0, sizeof(Q));
    if (k < 3 || k > MAXN) return 0;
    for (v = 0; v < n; v++)
        if (off[v + 1] - off[v] > maxd) maxd = (int32_t)(off[v + 1] - off[v]);
    for (d = 0; d < k; d++) Q.S[d] = (int32_t *)malloc(sizeof(int32_t) * (maxd + 1));
    Q.off = off;
    Q.adj = adj;
    Q.k = k;
    for (v0 = first; v0 < n; v0 += stride) {
        const int32_t *n0 = adj + off[v0];
        int64_t a0 = lower_bound_i32(n0, off[v0 + 1] - off[v0], v0), i;
        for (i = 0; i < a0; i++) { /* v1 in N+(v0) */
            int32_t v1 = n0[i];
            const int32_t *n1 = adj + off[v1];
            int64_t a = lower_bound_i32(n1, off[v1 + 1] - off[v1], v1);
            int64_t b = lower_bound_i32(n0, a0, v1), j, m = 0;
            for (j = 0; j < a; j++)
                if (in_sorted(n0, b, n1[j])) {
                    if (k > 3) Q.S[1][m] = n1[j];
                    m++;
                }
            /* father of the depth-1 op is the root neighbour list: +16 */
            if (k == 3) {
                Q.count += (uint64_t)m;
                Q.bytes += 4ull * (uint64_t)(a + b) + 32ull;
            } else {
                Q.bytes += 4ull * (uint64_t)(a + b) + 4ull * (uint64_t)m + 32ull;
                Q.Slen[1] = m;
                clq_rec(&Q, 2);
            }
        }
    }
    for (d = 0; d < k; d++) free(Q.S[d]);
    *walg = Q.bytes;
    return Q.count;
}

/* ------------------------------------------------------------------ */
static const int64_t *g_sort_off;
static int deg_cmp(const void *x, const void *y) {
    int32_t a = *(const int32_t *)x, b = *(const int32_t *)y;
    int64_t da = g_sort_off[a + 1] - g_sort_off[a], db = g_sort_off[b + 1] - g_sort_off[b];
    if (da != db) return da > db ? -1 : 1;
    return a < b ? -1 : (a > b);
}
static int i32_cmp(const void *x, const void *y) {
    int32_t a = *(const int32_t *)x, b = *(const int32_t *)y;
    return a < b ? -1 : (a > b);
}

void orc_relabel(int32_t n, const int64_t *off, const int32_t *adj, int64_t *off_out, int32_t *adj_out) {
    int32_t *order = (int32_t *)malloc(sizeof(int32_t) * (n + 1)), *nw = (int32_t *)malloc(sizeof(int32_t) * (n + 1));
    int32_t i;
    for (i = 0; i < n; i++) order[i] = i;
    g_sort_off = off;
    qsort(order, (size_t)n, sizeof(int32_t), deg_cmp);
    for (i = 0; i < n; i++) nw[order[i]] = i;
    off_out[0] = 0;
    for (i = 0; i < n; i++) {
        int32_t v = order[i];
        int64_t e, o = off_out[i];
        for (e = off[v]; e < off[v + 1]; e++) adj_out[o++] = nw[adj[e]];
        off_out[i + 1] = o;
        qsort(adj_out + off_out[i], (size_t)(o - off_out[i]), sizeof(int32_t), i32_cmp);
    }
    free(order);
    free(nw);
}

/* Algorithmic bytes of the bitmap clique store (cliques K3..kmax, kmax <= 6)
 * on a descending-degree graph. Per vertex v0 with d = |N+(v0)| >= 2 (L =
 * N+(v0)): 4d + 16 to read L; per position p of L, 4(|N+(L[p])| + p) + 16
 * to build row M_p = {q < p : L[q] in N+(L[p])} and 4 w(p) to count it,
 * where w(x) = floor(x/32) + 1 is the number of words that can hold bits
 * below x. For mutually adjacent positions p > q > r > s: 8 w(q) per pair
 * (p,q) (when kmax >= 4), 8 w(r) per triple (kmax >= 5) and 8 w(s) per
 * quadruple (kmax >= 6): the words of the two rows each step intersects.
 * Every clique with top vertex v0 is v0 plus a clique of these rows, so
 * #K3 = #pairs, #K4 = #triples, ... Returns #K_kmax. */
uint64_t orc_clique_bitmap_walg(int32_t n, const int64_t *off, const int32_t *adj, int kmax, uint64_t *walg) {
    return orc_clique_bitmap_walg_sample(n, off, adj, kmax, 1, 0, walg);
}

/* The same over the top vertices v0 = first, first + stride, ... only. */
uint64_t orc_clique_bitmap_walg_sample(int32_t n, const int64_t *off, const int32_t *adj, int kmax, int32_t stride,
                                       int32_t first, uint64_t *walg) {
    int32_t v0, maxd = 0, v;
    uint64_t bytes = 0, cnt[7] = {0};
    unsigned char *M; /* M[p*d+q] = 1 iff q < p and L[q] in N+(L[p]) */
    for (v = 0; v < n; v++)
        if (off[v + 1] - off[v] > maxd) maxd = (int32_t)(off[v + 1] - off[v]);
    M = (unsigned char *)malloc((size_t)(maxd + 1) * (maxd + 1));
    for (v0 = first; v0 < n; v0 += stride) {
        const int32_t *L = adj + off[v0];
        const int64_t d = lower_bound_i32(L, off[v0 + 1] - off[v0], v0);
        int64_t p, q, r, t, u;
        if (d < 2) continue;
        bytes += 4ull * (uint64_t)d + 16ull;
        for (p = 0; p < d; p++) {
            const int32_t x = L[p];
            const int64_t xl = lower_bound_i32(adj + off[x], off[x + 1] - off[x], x);
            bytes += 4ull * (uint64_t)(xl + p) + 16ull + 4ull * (uint64_t)(p / 32 + 1);
            for (q = 0; q < d; q++) M[p * d + q] = q < p && in_sorted(adj + off[x], xl, L[q]);
        }
        for (p = 0; p < d; p++)
            for (q = 0; q < p; q++) {
                if (!M[p * d + q]) continue;
                cnt[3]++;
                if (kmax < 4) continue;
                bytes += 8ull * (uint64_t)(q / 32 + 1);
                for (r = 0; r < q; r++) {
                    if (!(M[p * d + r] && M[q * d + r])) continue;
                    cnt[4]++;
                    if (kmax < 5) continue;
                    bytes += 8ull * (uint64_t)(r / 32 + 1);
                    for (t = 0; t < r; t++) {
                        if (!(M[p * d + t] && M[q * d + t] && M[r * d + t])) continue;
                        cnt[5]++;
                        if (kmax < 6) continue;
                        bytes += 8ull * (uint64_t)(t / 32 + 1);
                        for (u = 0; u < t; u++)
                            if (M[p * d + u] && M[q * d + u] && M[r * d + u] && M[t * d + u]) cnt[6]++;
                    }
                }
            }
    }
    free(M);
    *walg = bytes;
    return cnt[kmax];
}
