/*
 * oracle.h - CPU oracle for the push-style graph analytics path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load this library, and only as the checker or the timed CPU baseline;
 * the product (libeegraph.so) never calls it.
 *
 * Two kinds of functions:
 *   - restatements of the reference algorithm (Tigr layout, seq_cpu sweeps, delta-push PR),
 *     each citing the reference file:line it follows, with the reference defects D1-D6
 *     fixed as documented;
 *   - textbook oracles that pin the restatements (queue BFS, Dijkstra, max-min Dijkstra,
 *     union-find, FP64 Jacobi PageRank, Floyd-Warshall on tiny graphs).
 */
#ifndef EEG_ORACLE_H
#define EEG_ORACLE_H

#include <stdint.h>

#define ORC_INF 4294967294u

enum { ORC_BFS = 0, ORC_SSSP = 1, ORC_SSWP = 2, ORC_CC = 3 };

/* Stable counting sort of an edge list into out-edge CSR (row: n+1, col/wout: m). */
void orc_csr_build(uint32_t n, uint64_t m, const uint32_t* src, const uint32_t* dst,
                   const uint32_t* w, uint64_t* row, uint32_t* col, uint32_t* wout);

/* Tigr layout of src/virtual_graph.cpp:36-112 with D1 (truncation at num_edges) and D2
 * (uninitialised counter) fixed.  Sizes: node_pointer n, edge_list 2m+n, parts 2*P.
 * Returns P = number of parts.  Any output may be NULL (then only P is computed). */
uint64_t orc_tigr_layout(uint32_t n, uint64_t m, const uint32_t* src, const uint32_t* dst,
                         const uint32_t* w, uint32_t K, uint64_t* node_pointer,
                         uint32_t* edge_list, uint32_t* parts);

/* seq_cpu sweep (src/sssp.cu:357-404, src/bfs.cu:275-320, src/sswp.cu:273-320,
 * src/cc.cu:329-376) over a Tigr layout: Gauss-Seidel passes over parts in ascending order
 * until a pass changes nothing.  labels must hold the initial labels.  Relaxation only from
 * usable labels (fixes D3); CC uses min-label (fixes D4; pass the symmetrised layout).
 * time_cap_s > 0 stops early; *converged tells whether the fixpoint was reached.
 * Returns the number of passes run. */
uint32_t orc_seq_sweep(int alg, uint32_t n, uint64_t num_parts, const uint64_t* node_pointer,
                       const uint32_t* edge_list, const uint32_t* parts, uint32_t K,
                       uint32_t* labels, double time_cap_s, int* converged);

/* Initial labels as the reference drivers set them (src/eegraph.cu:38-49, 544-553,
 * 1549-1560). */
void orc_init_labels(int alg, uint32_t n, uint32_t source, uint32_t* labels);

/* Textbook oracles on a CSR. */
void orc_bfs(uint32_t n, const uint64_t* row, const uint32_t* col, uint32_t source, uint32_t* out);
void orc_dijkstra(uint32_t n, const uint64_t* row, const uint32_t* col, const uint32_t* w,
                  uint32_t source, uint32_t* out);
void orc_widest(uint32_t n, const uint64_t* row, const uint32_t* col, const uint32_t* w,
                uint32_t source, uint32_t* out);
/* The same answers as orc_dijkstra / orc_widest for integer weights in [0, wmax] with bucket
 * queues (Dial 1969), linear in m plus the largest label: for the largest graphs.  Return -1
 * (out untouched) if a weight exceeds wmax. */
int orc_dial(uint32_t n, const uint64_t* row, const uint32_t* col, const uint32_t* w, uint32_t wmax,
             uint32_t source, uint32_t* out);
int orc_dial_widest(uint32_t n, const uint64_t* row, const uint32_t* col, const uint32_t* w,
                    uint32_t wmax, uint32_t source, uint32_t* out);
/* Weakly connected components: label = minimum vertex id in the component. */
void orc_wcc(uint32_t n, uint64_t m, const uint32_t* src, const uint32_t* dst, uint32_t* out);

/* Floyd-Warshall style closures for n <= 64 (brute force pins for the oracles above):
 * alg selects shortest hops / shortest weighted / widest / weak connectivity. */
int orc_brute(int alg, uint32_t n, uint64_t m, const uint32_t* src, const uint32_t* dst,
              const uint32_t* w, uint32_t source, uint32_t* out);

/* FP64 Jacobi PageRank x = (1-d)/N + d * sum_{u->v} x_u/outdeg(u), x0 = 1/N (or a one-hot
 * restart at `source` with restart mass (1-d) when personalised != 0).  Stops when the L1
 * change drops below eps or after max_iter.  Returns iterations; *residual = last L1 change. */
uint32_t orc_pagerank(uint32_t n, const uint64_t* row, const uint32_t* col, double d, double eps,
                      uint32_t max_iter, int personalised, uint32_t source, double* x,
                      double* residual);

/* The same FP64 Jacobi iteration (also personalised) in pull form over the in-edge CSR
 * (in_row/in_col: sources of each vertex's in-edges; outdeg: out-degree of every vertex),
 * parallel over destinations with OpenMP (`threads` <= 0: the OpenMP default).  Each vertex's
 * sum is taken in in-edge order by one thread, so the ranks do not depend on the thread
 * count.  For large graphs, where orc_pagerank would take minutes. */
uint32_t orc_pagerank_pull(uint32_t n, const uint64_t* in_row, const uint32_t* in_col,
                           const uint32_t* outdeg, double d, double eps, uint32_t max_iter,
                           int threads, int personalised, uint32_t source, double* x,
                           double* residual);

/* Serial delta-push PageRank following src/pr.cu:171-225 (value += delta, push d*delta/deg)
 * with D6 fixed (each vertex consumes its delta once) and in the normalised scale
 * (x0 = 0, residual0 = (1-d)/N).  Stops when no vertex holds residual > thr.  Float state
 * like the reference.  Returns passes; time_cap_s > 0 stops early (then *converged = 0). */
uint32_t orc_pr_delta_push(uint32_t n, const uint64_t* row, const uint32_t* col, float d,
                           float thr, float* x, double time_cap_s, int* converged);

#endif
