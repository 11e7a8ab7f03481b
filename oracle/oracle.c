/*
 * oracle.c - CPU oracle for the push-style graph analytics path.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg as the checker / timed CPU baseline; never by the product.
 * See oracle.h for the contract of every function.
 */
#include "oracle.h"

#include <stdlib.h>
#include <math.h>
#include <omp.h>
#include <string.h>
#include <time.h>

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* Counting sort, stable by input order.  In parallel: every thread owns a range of sources
 * holding about m/T edges and scans the whole list in input order, placing only its own
 * sources' edges, so each row keeps input order whatever the thread count. */
void orc_csr_build(uint32_t n, uint64_t m, const uint32_t* src, const uint32_t* dst,
                   const uint32_t* w, uint64_t* row, uint32_t* col, uint32_t* wout) {
    memset(row, 0, sizeof(uint64_t) * ((size_t)n + 1));
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)m; ++i) {
#pragma omp atomic
        row[src[i] + 1]++;
    }
    for (uint32_t v = 0; v < n; ++v) row[v + 1] += row[v];
    uint64_t* cur = (uint64_t*)malloc(sizeof(uint64_t) * ((size_t)n + 1));
    memcpy(cur, row, sizeof(uint64_t) * ((size_t)n + 1));
#pragma omp parallel
    {
        const uint64_t T = (uint64_t)omp_get_num_threads(), t = (uint64_t)omp_get_thread_num();
        /* first source whose row starts at or after edge k (sources [lo, hi) are this thread's) */
#define ORC_FIRST_AT(k, out)                         \
        do {                                         \
            uint32_t a_ = 0, b_ = n;                 \
            while (a_ < b_) {                        \
                const uint32_t c_ = a_ + (b_ - a_) / 2; \
                if (row[c_] < (k)) a_ = c_ + 1; else b_ = c_; \
            }                                        \
            (out) = a_;                              \
        } while (0)
        uint32_t lo, hi;
        ORC_FIRST_AT(m * t / T, lo);
        ORC_FIRST_AT(m * (t + 1) / T, hi);
        if (t == 0) lo = 0;
        if (t + 1 == T) hi = n;
#undef ORC_FIRST_AT
        if (lo < hi) {
            for (uint64_t i = 0; i < m; ++i) {
                const uint32_t s = src[i];
                if (s < lo || s >= hi) continue;
                const uint64_t p = cur[s]++;
                col[p] = dst[i];
                if (wout) wout[p] = w ? w[i] : 1u;
            }
        }
    }
    free(cur);
}

/* src/virtual_graph.cpp:36-112.  nodePointer[v] = running counter advanced by 2*deg+1 (:50-64);
 * edgeList[nodePointer[v]] = outDegree (:53); edge number c of v (c = how many edges of v were
 * placed before, input order) at nodePointer+1+2c (dst) and +2+2c (weight) (:84-93);
 * parts (v, p) for p < ceil(deg/K) in ascending v (:97-111).  D1 fixed: no truncation test
 * against num_edges (:86).  D2 fixed: the per-vertex counter starts at zero (:66). */
uint64_t orc_tigr_layout(uint32_t n, uint64_t m, const uint32_t* src, const uint32_t* dst,
                         const uint32_t* w, uint32_t K, uint64_t* node_pointer,
                         uint32_t* edge_list, uint32_t* parts) {
    uint32_t* deg = (uint32_t*)calloc((size_t)n + 1, sizeof(uint32_t));
    for (uint64_t i = 0; i < m; ++i) deg[src[i]]++;  /* outDegree (src/virtual_graph.cpp:25-32) */
    uint64_t* np = (uint64_t*)malloc(sizeof(uint64_t) * ((size_t)n + 1));
    uint64_t counter = 0, P = 0;
    for (uint32_t v = 0; v < n; ++v) {
        np[v] = counter;
        if (edge_list) edge_list[counter] = deg[v];
        P += (deg[v] + K - 1) / K;
        counter += 2ull * deg[v] + 1;
    }
    if (node_pointer) memcpy(node_pointer, np, sizeof(uint64_t) * n);
    if (edge_list) {
        uint32_t* placed = (uint32_t*)calloc((size_t)n + 1, sizeof(uint32_t));
        for (uint64_t i = 0; i < m; ++i) {
            const uint32_t s = src[i];
            const uint64_t loc = np[s] + 1 + 2ull * placed[s];
            edge_list[loc] = dst[i];
            edge_list[loc + 1] = w ? w[i] : 1u;
            placed[s]++;
        }
        free(placed);
    }
    if (parts) {
        uint64_t c = 0;
        for (uint32_t v = 0; v < n; ++v) {
            const uint32_t np_v = (deg[v] + K - 1) / K;
            for (uint32_t p = 0; p < np_v; ++p) {
                parts[2 * c] = v;
                parts[2 * c + 1] = p;
                ++c;
            }
        }
    }
    free(np);
    free(deg);
    return P;
}

void orc_init_labels(int alg, uint32_t n, uint32_t source, uint32_t* labels) {
    for (uint32_t v = 0; v < n; ++v) {
        switch (alg) {
            case ORC_BFS:
            case ORC_SSSP: labels[v] = ORC_INF; break;  /* src/eegraph.cu:38-49 */
            case ORC_SSWP: labels[v] = 0; break;        /* src/eegraph.cu:1549-1560 */
            default: labels[v] = v; break;              /* src/eegraph.cu:544-547 */
        }
    }
    if (alg == ORC_BFS || alg == ORC_SSSP) labels[source] = 0;
    if (alg == ORC_SSWP) labels[source] = ORC_INF;
}

/* The relaxation of one part, as in src/sssp.cu:369-401: the source label is read once per
 * part, edge i of the part is j = part + i*numParts(v) (strided), stored at
 * nodePointer + 1 + 2j (dst) and +2+2j (weight).  BFS uses +1 (src/bfs.cu:309), SSWP
 * min(label, w) with max (src/sswp.cu:309-312), CC the plain label with min. */
uint32_t orc_seq_sweep(int alg, uint32_t n, uint64_t num_parts, const uint64_t* node_pointer,
                       const uint32_t* edge_list, const uint32_t* parts, uint32_t K,
                       uint32_t* labels, double time_cap_s, int* converged) {
    (void)n;
    const double t0 = now_s();
    uint32_t passes = 0;
    int finished;
    *converged = 0;
    do {
        finished = 1;
        ++passes;
        for (uint64_t pi = 0; pi < num_parts; ++pi) {
            const uint32_t v = parts[2 * pi], part = parts[2 * pi + 1];
            const uint32_t lab = labels[v];
            const uint64_t np = node_pointer[v];
            const uint32_t deg = edge_list[np];
            const uint32_t nparts = (deg + K - 1) / K;
            if ((alg == ORC_BFS || alg == ORC_SSSP) && lab >= ORC_INF) continue; /* D3 fix */
            if (alg == ORC_SSWP && lab == 0) continue;
            for (uint32_t i = 0; i < K; ++i) {
                const uint64_t j = (uint64_t)part + (uint64_t)i * nparts;
                if (j >= deg) break;
                const uint64_t loc = np + 1 + 2 * j;
                const uint32_t d = edge_list[loc], w = edge_list[loc + 1];
                uint32_t c;
                switch (alg) {
                    case ORC_BFS: c = lab + 1; break;
                    case ORC_SSSP: {
                        const uint64_t s = (uint64_t)lab + w;
                        c = s < ORC_INF ? (uint32_t)s : ORC_INF;
                        break;
                    }
                    case ORC_SSWP: c = lab < w ? lab : w; break;
                    default: c = lab; break;
                }
                const int better = (alg == ORC_SSWP) ? (c > labels[d]) : (c < labels[d]);
                if (better) {
                    labels[d] = c;
                    finished = 0;
                }
            }
        }
        if (!finished && time_cap_s > 0 && now_s() - t0 > time_cap_s) return passes;
    } while (!finished);
    *converged = 1;
    return passes;
}

/* ---------------------------------------------------------------- textbook oracles */

void orc_bfs(uint32_t n, const uint64_t* row, const uint32_t* col, uint32_t source, uint32_t* out) {
    for (uint32_t v = 0; v < n; ++v) out[v] = ORC_INF;
    uint32_t* q = (uint32_t*)malloc(sizeof(uint32_t) * ((size_t)n + 1));
    uint64_t head = 0, tail = 0;
    out[source] = 0;
    q[tail++] = source;
    while (head < tail) {
        const uint32_t u = q[head++];
        for (uint64_t e = row[u]; e < row[u + 1]; ++e) {
            const uint32_t v = col[e];
            if (out[v] == ORC_INF) {
                out[v] = out[u] + 1;
                q[tail++] = v;
            }
        }
    }
    free(q);
}

/* binary heap of (key, vertex) with lazy deletion; min-heap if sign = +1, max-heap if -1 */
typedef struct {
    uint64_t key;
    uint32_t v;
} HeapItem;
typedef struct {
    HeapItem* a;
    uint64_t n, cap;
    int maxheap;
} Heap;
static int heap_before(const Heap* h, HeapItem x, HeapItem y) {
    return h->maxheap ? (x.key > y.key) : (x.key < y.key);
}
static void heap_push(Heap* h, uint64_t key, uint32_t v) {
    if (h->n == h->cap) {
        h->cap = h->cap ? 2 * h->cap : 1024;
        h->a = (HeapItem*)realloc(h->a, sizeof(HeapItem) * h->cap);
    }
    uint64_t i = h->n++;
    HeapItem x = {key, v};
    while (i > 0) {
        const uint64_t p = (i - 1) / 2;
        if (!heap_before(h, x, h->a[p])) break;
        h->a[i] = h->a[p];
        i = p;
    }
    h->a[i] = x;
}
static HeapItem heap_pop(Heap* h) {
    HeapItem top = h->a[0], x = h->a[--h->n];
    uint64_t i = 0;
    for (;;) {
        uint64_t c = 2 * i + 1;
        if (c >= h->n) break;
        if (c + 1 < h->n && heap_before(h, h->a[c + 1], h->a[c])) ++c;
        if (!heap_before(h, h->a[c], x)) break;
        h->a[i] = h->a[c];
        i = c;
    }
    if (h->n) h->a[i] = x;
    return top;
}

void orc_dijkstra(uint32_t n, const uint64_t* row, const uint32_t* col, const uint32_t* w,
                  uint32_t source, uint32_t* out) {
    uint64_t* dist = (uint64_t*)malloc(sizeof(uint64_t) * ((size_t)n + 1));
    for (uint32_t v = 0; v < n; ++v) dist[v] = UINT64_MAX;
    Heap h = {0, 0, 0, 0};
    dist[source] = 0;
    heap_push(&h, 0, source);
    while (h.n) {
        HeapItem it = heap_pop(&h);
        if (it.key != dist[it.v]) continue;
        for (uint64_t e = row[it.v]; e < row[it.v + 1]; ++e) {
            const uint64_t nd = it.key + (w ? w[e] : 1u);
            if (nd < dist[col[e]]) {
                dist[col[e]] = nd;
                heap_push(&h, nd, col[e]);
            }
        }
    }
    for (uint32_t v = 0; v < n; ++v) out[v] = dist[v] >= ORC_INF ? ORC_INF : (uint32_t)dist[v];
    free(h.a);
    free(dist);
}

void orc_widest(uint32_t n, const uint64_t* row, const uint32_t* col, const uint32_t* w,
                uint32_t source, uint32_t* out) {
    for (uint32_t v = 0; v < n; ++v) out[v] = 0;
    Heap h = {0, 0, 0, 1};
    out[source] = ORC_INF;
    heap_push(&h, ORC_INF, source);
    while (h.n) {
        HeapItem it = heap_pop(&h);
        if (it.key != out[it.v]) continue;
        for (uint64_t e = row[it.v]; e < row[it.v + 1]; ++e) {
            const uint32_t we = w ? w[e] : 1u;
            const uint32_t c = (uint32_t)it.key < we ? (uint32_t)it.key : we;
            if (c > out[col[e]]) {
                out[col[e]] = c;
                heap_push(&h, c, col[e]);
            }
        }
    }
    free(h.a);
}

/* Growable vertex list of one bucket. */
typedef struct {
    uint32_t* a;
    size_t n, cap;
} Bucket;
static void bucket_push(Bucket* b, uint32_t v) {
    if (b->n == b->cap) {
        b->cap = b->cap ? 2 * b->cap : 64;
        b->a = (uint32_t*)realloc(b->a, sizeof(uint32_t) * b->cap);
    }
    b->a[b->n++] = v;
}

/* Dial's bucket queue (R. B. Dial, "Algorithm 360: shortest-path forest with topological
 * ordering", CACM 12(11), 1969) for integer weights in [0, wmax]: wmax+1 circular buckets
 * indexed by distance mod (wmax+1); the smallest tentative distance is the current bucket, so
 * each vertex is settled when its bucket comes round.  Same answer as orc_dijkstra in
 * O(m + max distance) instead of O(m log n). */
int orc_dial(uint32_t n, const uint64_t* row, const uint32_t* col, const uint32_t* w, uint32_t wmax,
             uint32_t source, uint32_t* out) {
    if (w) {
        for (uint64_t e = 0; e < row[n]; ++e)
            if (w[e] > wmax) return -1;
    }
    const size_t nb = (size_t)wmax + 1;
    Bucket* bk = (Bucket*)calloc(nb, sizeof(Bucket));
    uint64_t* dist = (uint64_t*)malloc(sizeof(uint64_t) * ((size_t)n + 1));
    for (uint32_t v = 0; v < n; ++v) dist[v] = UINT64_MAX;
    dist[source] = 0;
    bucket_push(&bk[0], source);
    uint64_t queued = 1, cur = 0;
    while (queued) {
        Bucket* b = &bk[cur % nb];
        for (size_t i = 0; i < b->n; ++i) {  /* weight-0 edges append to this bucket */
            const uint32_t v = b->a[i];
            --queued;
            if (dist[v] != cur) continue;  /* settled earlier at a smaller distance */
            for (uint64_t e = row[v]; e < row[v + 1]; ++e) {
                const uint64_t nd = cur + (w ? w[e] : 1u);
                if (nd < dist[col[e]]) {
                    dist[col[e]] = nd;
                    bucket_push(&bk[nd % nb], col[e]);
                    ++queued;
                }
            }
        }
        b->n = 0;
        ++cur;
    }
    for (uint32_t v = 0; v < n; ++v) out[v] = dist[v] >= ORC_INF ? ORC_INF : (uint32_t)dist[v];
    for (size_t i = 0; i < nb; ++i) free(bk[i].a);
    free(bk);
    free(dist);
    return 0;
}

/* Widest (max-min) paths with a bucket per label value, for integer weights in [0, wmax]: a
 * label only moves down along a path, so buckets are settled from wmax down to 1; the source
 * holds ORC_INF and gives each out-neighbour its edge weight.  Same answer as orc_widest. */
int orc_dial_widest(uint32_t n, const uint64_t* row, const uint32_t* col, const uint32_t* w,
                    uint32_t wmax, uint32_t source, uint32_t* out) {
    if (w) {
        for (uint64_t e = 0; e < row[n]; ++e)
            if (w[e] > wmax) return -1;
    }
    const size_t nb = (size_t)wmax + 1;
    Bucket* bk = (Bucket*)calloc(nb, sizeof(Bucket));
    for (uint32_t v = 0; v < n; ++v) out[v] = 0;
    out[source] = ORC_INF;
    for (uint64_t e = row[source]; e < row[source + 1]; ++e) {
        const uint32_t c = w ? w[e] : 1u;
        if (c > out[col[e]]) {
            out[col[e]] = c;
            bucket_push(&bk[c], col[e]);
        }
    }
    for (size_t lv = wmax; lv >= 1; --lv) {
        Bucket* b = &bk[lv];
        for (size_t i = 0; i < b->n; ++i) {  /* edges of weight >= lv append to this bucket */
            const uint32_t v = b->a[i];
            if (out[v] != lv) continue;  /* raised to a higher bucket meanwhile */
            for (uint64_t e = row[v]; e < row[v + 1]; ++e) {
                const uint32_t we = w ? w[e] : 1u;
                const uint32_t c = we < lv ? we : (uint32_t)lv;
                if (c > out[col[e]]) {
                    out[col[e]] = c;
                    bucket_push(&bk[c], col[e]);
                }
            }
        }
        b->n = 0;
    }
    for (size_t i = 0; i < nb; ++i) free(bk[i].a);
    free(bk);
    return 0;
}

static uint32_t uf_find(uint32_t* p, uint32_t x) {
    while (p[x] != x) {
        p[x] = p[p[x]];
        x = p[x];
    }
    return x;
}

void orc_wcc(uint32_t n, uint64_t m, const uint32_t* src, const uint32_t* dst, uint32_t* out) {
    uint32_t* p = (uint32_t*)malloc(sizeof(uint32_t) * ((size_t)n + 1));
    for (uint32_t v = 0; v < n; ++v) p[v] = v;
    for (uint64_t i = 0; i < m; ++i) {
        uint32_t a = uf_find(p, src[i]), b = uf_find(p, dst[i]);
        if (a == b) continue;
        if (a < b) p[b] = a; else p[a] = b;  /* the root is always the smaller id */
    }
    for (uint32_t v = 0; v < n; ++v) out[v] = uf_find(p, v);
    free(p);
}

int orc_brute(int alg, uint32_t n, uint64_t m, const uint32_t* src, const uint32_t* dst,
              const uint32_t* w, uint32_t source, uint32_t* out) {
    if (n == 0 || n > 64 || source >= n) return -1;
    uint64_t d[64][64];
    const uint64_t BIG = UINT64_MAX / 4;
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t j = 0; j < n; ++j) {
            if (alg == ORC_SSWP) d[i][j] = 0;
            else if (alg == ORC_CC) d[i][j] = (i == j);
            else d[i][j] = (i == j) ? 0 : BIG;
        }
    for (uint64_t e = 0; e < m; ++e) {
        const uint32_t a = src[e], b = dst[e];
        const uint64_t we = (alg == ORC_BFS) ? 1 : (w ? w[e] : 1u);
        if (alg == ORC_SSWP) { if (we > d[a][b]) d[a][b] = we; }
        else if (alg == ORC_CC) { d[a][b] = 1; d[b][a] = 1; }
        else if (we < d[a][b]) d[a][b] = we;
    }
    for (uint32_t k = 0; k < n; ++k)
        for (uint32_t i = 0; i < n; ++i)
            for (uint32_t j = 0; j < n; ++j) {
                if (alg == ORC_SSWP) {
                    const uint64_t c = d[i][k] < d[k][j] ? d[i][k] : d[k][j];
                    if (c > d[i][j]) d[i][j] = c;
                } else if (alg == ORC_CC) {
                    if (d[i][k] && d[k][j]) d[i][j] = 1;
                } else if (d[i][k] + d[k][j] < d[i][j]) {
                    d[i][j] = d[i][k] + d[k][j];
                }
            }
    for (uint32_t v = 0; v < n; ++v) {
        if (alg == ORC_CC) {
            uint32_t mn = v;
            for (uint32_t j = 0; j < n; ++j)
                if (d[v][j] && j < mn) mn = j;
            out[v] = mn;
        } else if (alg == ORC_SSWP) {
            const uint64_t x = (v == source) ? ORC_INF : d[source][v];
            out[v] = x >= ORC_INF ? ORC_INF : (uint32_t)x;
        } else {
            const uint64_t x = d[source][v];
            out[v] = x >= ORC_INF ? ORC_INF : (uint32_t)x;
        }
    }
    return 0;
}

uint32_t orc_pagerank(uint32_t n, const uint64_t* row, const uint32_t* col, double d, double eps,
                      uint32_t max_iter, int personalised, uint32_t source, double* x,
                      double* residual) {
    double* nx = (double*)malloc(sizeof(double) * ((size_t)n + 1));
    for (uint32_t v = 0; v < n; ++v) x[v] = personalised ? (v == source ? 1.0 : 0.0) : 1.0 / n;
    const double base = (1.0 - d) / n;
    uint32_t it = 0;
    double res = 0;
    while (it < max_iter) {
        for (uint32_t v = 0; v < n; ++v) nx[v] = personalised ? 0.0 : base;
        if (personalised) nx[source] = 1.0 - d;
        for (uint32_t u = 0; u < n; ++u) {
            const uint64_t deg = row[u + 1] - row[u];
            if (!deg) continue;
            const double c = d * x[u] / (double)deg;
            for (uint64_t e = row[u]; e < row[u + 1]; ++e) nx[col[e]] += c;
        }
        res = 0;
        for (uint32_t v = 0; v < n; ++v) {
            const double df = nx[v] - x[v];
            res += df < 0 ? -df : df;
            x[v] = nx[v];
        }
        ++it;
        if (res < eps) break;
    }
    if (residual) *residual = res;
    free(nx);
    return it;
}

uint32_t orc_pagerank_pull(uint32_t n, const uint64_t* in_row, const uint32_t* in_col,
                           const uint32_t* outdeg, double d, double eps, uint32_t max_iter,
                           int threads, int personalised, uint32_t source, double* x,
                           double* residual) {
    double* nx = (double*)malloc(sizeof(double) * ((size_t)n + 1));
    double* c = (double*)malloc(sizeof(double) * ((size_t)n + 1));
    if (threads > 0) omp_set_num_threads(threads);
    const double base = personalised ? 0.0 : (1.0 - d) / n;
    for (uint32_t v = 0; v < n; ++v) x[v] = personalised ? (v == source ? 1.0 : 0.0) : 1.0 / n;
    uint32_t it = 0;
    double res = 0;
    while (it < max_iter) {
#pragma omp parallel for schedule(static)
        for (int64_t v = 0; v < (int64_t)n; ++v) c[v] = outdeg[v] ? d * x[v] / (double)outdeg[v] : 0.0;
        res = 0;
#pragma omp parallel for schedule(dynamic, 4096) reduction(+ : res)
        for (int64_t v = 0; v < (int64_t)n; ++v) {
            double s = (personalised && (uint32_t)v == source) ? 1.0 - d : base;
            for (uint64_t e = in_row[v]; e < in_row[v + 1]; ++e) s += c[in_col[e]];
            nx[v] = s;
            res += fabs(s - x[v]);
        }
        memcpy(x, nx, sizeof(double) * n);
        ++it;
        if (res < eps) break;
    }
    if (residual) *residual = res;
    free(nx);
    free(c);
    return it;
}

/* src/pr.cu:171-225 (async_push_td) restated serially per vertex: a vertex holding residual
 * above the threshold adds it to its value (:193-195) and pushes d*r/deg along every
 * out-edge (:208-217), then subtracts what it consumed (:220).  The reference repeats this
 * per part of a vertex (D6); here each vertex consumes its residual once per pass. */
uint32_t orc_pr_delta_push(uint32_t n, const uint64_t* row, const uint32_t* col, float d,
                           float thr, float* x, double time_cap_s, int* converged) {
    float* r = (float*)malloc(sizeof(float) * ((size_t)n + 1));
    const float r0 = (1.0f - d) / (float)n;
    for (uint32_t v = 0; v < n; ++v) {
        x[v] = 0.0f;
        r[v] = r0;
    }
    const double t0 = now_s();
    uint32_t passes = 0;
    int finished;
    *converged = 0;
    do {
        finished = 1;
        ++passes;
        for (uint32_t v = 0; v < n; ++v) {
            const float take = r[v];
            if (!(take > thr)) continue;
            x[v] += take;
            const uint64_t deg = row[v + 1] - row[v];
            if (deg) {
                finished = 0;
                const float push = (take / (float)deg) * d;
                for (uint64_t e = row[v]; e < row[v + 1]; ++e) r[col[e]] += push;
            }
            r[v] -= take;
        }
        if (!finished && time_cap_s > 0 && now_s() - t0 > time_cap_s) {
            free(r);
            return passes;
        }
    } while (!finished);
    *converged = 1;
    free(r);
    return passes;
}
