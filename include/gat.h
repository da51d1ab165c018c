/*
 * gat.h -- C ABI of the MI355X subgraph pattern-matching engine (libgat.so).
 *
 * The reference has no FFI: its boundary is an in-process C++ template API
 * plus a process/file protocol. Each entry point below names the reference
 * interface it replaces (paths relative to the reference repository root).
 *
 * Conventions
 *   - Every function returns 0 (GAT_OK) on success and a negative GAT_E*
 *     code on failure; the message is available from gat_last_error() on the
 *     same thread. Nothing here calls exit(); the reference's gpuErrchk and
 *     graph-mismatch paths exit the process instead
 *     (include/utils/cuda_utils.cuh:9-16, include/infra/graph_backend.cuh:41-45).
 *   - Vertex ids are int32, CSR offsets int64 (include/core/types.hpp:4-5).
 *     Counts are uint64 (the reference uses long long, target/pm.cu:62).
 *   - A handle is used by one host thread at a time. Each gat_plan owns one
 *     HIP stream; the GPU entry points need a visible gfx950 device and fail
 *     with GAT_ENODEV otherwise (there is no CPU fallback).
 */
#ifndef GAT_H
#define GAT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define GAT_OK 0
#define GAT_EINVAL (-1)      /* bad argument / malformed pattern or graph  */
#define GAT_EIO (-2)         /* file could not be read or written          */
#define GAT_ENOMEM (-3)      /* host or device allocation failed           */
#define GAT_ENODEV (-4)      /* no usable GPU / HIP runtime error          */
#define GAT_ELIMIT (-5)      /* a fixed capacity would be exceeded (also a
                                device work list that filled up)           */
#define GAT_EUNSUPPORTED (-6)/* the request has no defined result          */
#define GAT_EINFEASIBLE (-7) /* kernel config cannot launch (LDS, block)   */
#define GAT_EINTERNAL (-8)   /* the engine flagged an inconsistency of its own
                                (a bug, never a property of the input)     */

/* ---- capacities of the schedule record ------------------------------ */
/* MAX_VERTEXES / MAX_PREFIXS of include/consts/general_consts.hpp:3-4 are
 * 10 / 10; the prefix capacity is raised to 16 because some 6-vertex
 * schedules need 11-12 prefixes (the reference overflows silently). */
#define GAT_MAX_VERTEXES 10
#define GAT_MAX_PREFIXS 16
#define GAT_MAX_IEP_GROUPS 64 /* schedule.hpp:962 has 50 */

/* Counting semantics.
 *   GAT_MODE_UNIQUE    : number of distinct non-induced subgraphs of the
 *                        data graph isomorphic to the pattern.
 *   GAT_MODE_REFERENCE : the value the reference engine computes for the
 *                        same schedule (include/engine/engine_kernel.cuh,
 *                        include/core/schedule.hpp:828-831, :955); k*#K_k
 *                        for a k-clique.
 *   GAT_MODE_LABELED   : labeled embeddings (|Aut| * unique), the quantity
 *                        target/naive_pm.cu:10-44 counts. */
#define GAT_MODE_UNIQUE 0
#define GAT_MODE_REFERENCE 1
#define GAT_MODE_LABELED 2

/* How a plan picks each pattern's Schedule. */
#define GAT_SCHEDULE_REFERENCE 0 /* the reference's pick: cost estimate under its
                                    fixed stats (5000, 1e6, 1e4), schedule.hpp:909-910 */
#define GAT_SCHEDULE_GRAPH 1     /* unique/labeled only: candidates ranked by the
                                    same estimate under this graph's vertex,
                                    directed-edge and triangle counts */
#define GAT_SCHEDULE_RANKED 2    /* unique/labeled only: candidates ranked under
                                    the reference's fixed statistics, so a rank
                                    names the same Schedule on every graph
                                    (rank 0 = the reference's pick) */

/* ---- opaque handles -------------------------------------------------- */
typedef struct gat_graph gat_graph; /* host CSR (+ device copy once used) */
typedef struct gat_plan gat_plan;   /* compiled schedules + device state   */

/* Flat, pointer-free schedule record. Field-for-field superset of
 * Core::ScheduleData (include/core/schedule.hpp:984-1026); the extra fields
 * carry what the engine needs beyond the reference's record (the chosen
 * permutation, the unfiltered restriction list, the automorphism count). */
typedef struct gat_schedule_data {
    int32_t v_cnt;                  /* pattern vertices                   */
    int32_t iep_suffix_vertexes;    /* schedule.hpp:595                   */
    int32_t basic_vertexes;         /* schedule.hpp:597                   */
    int32_t total_prefix_num;
    int32_t total_restrictions_num; /* after the second<basic filter      */
    int32_t prefix_depth[GAT_MAX_PREFIXS];
    int32_t prefix_data[GAT_MAX_PREFIXS][GAT_MAX_VERTEXES];
    int32_t prefixs_father[GAT_MAX_PREFIXS];
    int32_t prefixs_size_only[GAT_MAX_PREFIXS];
    int32_t vertex_prefix_start[GAT_MAX_VERTEXES + 1];
    int32_t loop_set_prefix_id[GAT_MAX_VERTEXES];
    int32_t restrictions_start[GAT_MAX_VERTEXES + 1];
    int32_t restrictions[GAT_MAX_VERTEXES * GAT_MAX_VERTEXES];        /* first  */
    int32_t restriction_second[GAT_MAX_VERTEXES * GAT_MAX_VERTEXES]; /* second */
    int32_t iep_prefix_num;
    int32_t subgroups_num;
    int32_t iep_vertex_id[GAT_MAX_PREFIXS];
    int32_t iep_ans_pos[GAT_MAX_IEP_GROUPS];
    int32_t iep_coef[GAT_MAX_IEP_GROUPS];
    int32_t iep_flag[GAT_MAX_IEP_GROUPS];
    /* extras */
    int32_t perm[GAT_MAX_VERTEXES];  /* input vertex i -> schedule vertex perm[i] */
    int32_t n_all_restrictions;      /* best restriction set, unfiltered  */
    int32_t all_restr_first[GAT_MAX_VERTEXES * GAT_MAX_VERTEXES];
    int32_t all_restr_second[GAT_MAX_VERTEXES * GAT_MAX_VERTEXES];
    int32_t automorphisms;           /* |Aut(P)|                          */
    uint32_t adj_rows[GAT_MAX_VERTEXES]; /* permuted pattern, bit j of row i */
    double best_cost;                /* estimate_restrictions of the pick */
} gat_schedule_data;

/* Runtime kernel configuration. The reference fixes these at compile time
 * (include/configs/engine_config.hpp, vertex_set_config.hpp,
 * include/generated/default_config.hpp:5); here they
 * are runtime values, and variants that must be compile-time are
 * pre-instantiated. Zero in any field selects the engine default. */
typedef struct gat_kernel_config {
    int32_t block_waves;      /* waves per workgroup (threads_per_block/64)   */
    int32_t waves_per_cu;     /* persistent grid = waves_per_cu * CUs (0: 16
                                 for the array store, 8 for the clique store) */
    int32_t queue_chunk;      /* directed edges taken per work-queue grab      */
    int32_t slot_capacity;    /* int32 entries per LDS prefix-set slot         */
    int32_t leaf_mode;        /* 0 flattened, else lanes per candidate 1..64
                                 (threads_per_warp, engine_config.hpp:6)      */
    int32_t intersect_strategy; /* how one wave intersects two sorted lists:
                                 0 auto: probe the side that sits in LDS or
                                   has a hub bitmap, else stream the shorter
                                   list (binary search in the other);
                                 1 stream the new vertex's list, search the
                                   other;
                                 2 stream-shorter: stream the shorter list,
                                   binary search in the longer;
                                 3 merge path: the lanes split the merged
                                   order of both lists into equal runs;
                                 4 galloping: lanes walk blocks of the shorter
                                   list with exponential searches from their
                                   previous position in the longer.
                                 3 and 4 apply to the wave-wide intersections
                                 of non-leaf levels; leaf levels (a flattened
                                 sweep over many candidates) use 0 there   */
    int32_t gallop_ratio;     /* auto: stream the side that is not in LDS or a
                                 hub only when it is at most this many times
                                 longer than that side                       */
    int32_t search_linear_threshold; /* set_search_type: linear scan below this
                                 size, binary above (array_vertex_set.cuh:145) */
    int32_t max_regs;         /* register cap hint: 0, 64 or 128 VGPRs         */
    int32_t vertex_store;     /* 0 auto, 1 sorted arrays only, 2 bitmap rows for
                                 clique plans (VertexStoreType {Array, Bitmap},
                                 vertex_set_config.hpp:3-6)                    */
    int32_t hub_degree;       /* array store: vertices of at least this degree
                                 get a dense neighbourhood bitmap in HBM that
                                 replaces binary search in their lists
                                 (0 = 256, negative = none)                   */
    int32_t split_us;         /* array store: a task still running after this
                                 many microseconds, once its launch's queue
                                 is drained, hands the unfinished rest of its
                                 shallowest loop at depth >= 2 to subtask
                                 launches in up to 64 parts (0 = 2000,
                                 negative = never); counts and W_alg do not
                                 change                                       */
} gat_kernel_config;

/* Per-call statistics filled by gat_count (any pointer may be NULL). */
typedef struct gat_count_stats {
    double seconds;           /* device time of the count (hipEvents)        */
    double main_kernel_ms;    /* duration of the main engine launch          */
    double second_kernel_ms;  /* array store: the split and HBM overflow
                                 launches after the main one; bitmap store:
                                 workgroup tier 0                             */
    uint64_t overflow_tasks;  /* array store: tasks re-run with HBM-resident
                                 prefix sets; bitmap store: vertices passed to
                                 the workgroup tiers                          */
    uint64_t alg_bytes;       /* W_alg of the whole count (GAT_FLAG_ACCOUNT)  */
    uint64_t alg_bytes_main;  /* W_alg of the main launch alone              */
    uint64_t alg_bytes_second;/* W_alg of the launches second_kernel_ms times  */
    uint64_t tasks;           /* directed edges of the roots counted         */
    int32_t grid_blocks;      /* launched workgroups                         */
    int32_t block_threads;
    int32_t lds_bytes;        /* dynamic LDS per workgroup                   */
    int32_t split_tasks;      /* array store: subtasks handed to the split
                                 launch                                      */
    int32_t split_overflow;   /* of those, re-run by the HBM launch          */
    int64_t split_level[4];   /* array store: subtasks made for LDS levels 1,
                                 2 and 3 and for the HBM list (which also
                                 takes the split_overflow copies)             */
} gat_count_stats;

#define GAT_FLAG_ACCOUNT 1u   /* instrumented pass: accumulate W_alg bytes  */
#define GAT_FLAG_TIMELINE 2u  /* record per-wave / per-workgroup start and end
                                 times of every launch (gat_plan_timeline)  */

/* Plan options (gat_plan_create_ex). Zero-initialised = gat_plan_create. */
typedef struct gat_plan_options {
    int32_t schedule_policy;  /* GAT_SCHEDULE_*                              */
    int32_t schedule_rank;    /* GRAPH: take the rank-th candidate (0 = best),
                                 clamped to the last one                      */
    const int32_t *schedule_ranks; /* GRAPH: one rank per pattern (overrides
                                 schedule_rank), or NULL                      */
    int64_t reserved[2];
} gat_plan_options;

/* ---- errors ---------------------------------------------------------- */
const char *gat_last_error(void);
const char *gat_version(void);

/* ---- patterns and schedules ----------------------------------------- */
/* Replaces Core::Pattern{adj_str} + Core::Schedule{pattern}
 * (include/core/pattern.hpp:22-39, include/core/schedule.hpp:864-958) and
 * the ScheduleData record (schedule.hpp:984-1026). `out` must point to a
 * gat_schedule_data; `cap` is its size in bytes. */
int gat_schedule_build(const char *adj_str, gat_schedule_data *out, size_t cap);

/* Every (permutation, restriction set) Schedule candidate of the reference's
 * search at the largest IEP suffix (schedule.hpp:878-926), ranked by
 * estimate_restrictions (schedule.hpp:362-437) under the given graph
 * statistics, stably (with v_cnt=5000, e_cnt=1e6, tri_cnt=1e4 entry 0 is
 * gat_schedule_build's schedule), without candidates that compile to the same
 * plan. Writes up to `cap` records of `elem_size` bytes each to `out` and the
 * total number of candidates to *n_out. */
int gat_schedule_candidates(const char *adj_str, int32_t v_cnt, int64_t e_cnt, int64_t tri_cnt,
                            gat_schedule_data *out, size_t elem_size, int32_t cap, int32_t *n_out);

/* |Aut(P)| of a pattern string (schedule.hpp:189-191). */
int gat_pattern_automorphisms(const char *adj_str, int64_t *out);

/* ---- graphs ---------------------------------------------------------- */
/* Replaces Infra::GlobalMemoryGraph(file, binary=true)
 * (include/infra/graph_backend.cuh:91-99,155-161): int32 v_cnt, int64 e_cnt
 * (directed entries), int64 offsets[v_cnt+1], int32 adj[e_cnt]. */
int gat_graph_load_bin(const char *path, gat_graph **out);
/* Text form of graph_backend.cuh:23-89: "v_cnt e_cnt" then e_cnt lines
 * "u v"; ids relabelled by sorted value, self-loops dropped, symmetrised,
 * sorted, deduplicated. */
int gat_graph_load_txt(const char *path, gat_graph **out);
/* graph_backend.cuh:163-169 */
int gat_graph_save_bin(const gat_graph *g, const char *path);
/* Copy an existing CSR (offsets[v_cnt+1], adj[e_directed]); the adjacency
 * rows must be sorted, free of duplicates and self-loops, and symmetric. */
int gat_graph_from_csr(int32_t v_cnt, int64_t e_directed, const int64_t *offsets,
                       const int32_t *adj, gat_graph **out);
/* Build from an undirected edge list (any order, duplicates and self-loops
 * allowed): symmetrised, sorted, deduplicated. */
int gat_graph_from_edges(int32_t v_cnt, int64_t n_edges, const int32_t *src,
                         const int32_t *dst, gat_graph **out);
/* Seeded generators (splitmix64-seeded xoshiro256**). utils/random_graph.cpp
 * draws from unseeded rand(); these replace it with reproducible streams. */
int gat_graph_gen_er(int32_t n, double p, uint64_t seed, gat_graph **out);
int gat_graph_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c,
                       uint64_t seed, gat_graph **out);
int gat_graph_gen_chung_lu(int32_t n, double mean_degree, double gamma, int32_t max_degree,
                           uint64_t seed, gat_graph **out);
int gat_graph_info(const gat_graph *g, int32_t *v_cnt, int64_t *e_directed,
                   int32_t *max_degree);
/* Read-only views of the host CSR (valid until gat_graph_free). */
int gat_graph_csr(const gat_graph *g, const int64_t **offsets, const int32_t **adj);
/* 32-bin log2-degree histogram (bin k counts vertices with
 * floor(log2(deg+1)) == k), used as cost-model features. */
int gat_graph_degree_histogram(const gat_graph *g, int64_t *bins32);
/* Number of triangles (host count on the degree-ordered graph, cached). */
int gat_graph_triangles(gat_graph *g, int64_t *out);
void gat_graph_free(gat_graph *g);

/* ---- plans and counting --------------------------------------------- */
/* Replaces Engine::DeviceContext{schedule, graph} + context.to_device() +
 * Engine::Executor{set_size, GPU_DEVICE} (include/engine/context.cuh:18-25,
 * include/engine/engine.cuh:157-204) for n_patterns schedules at once.
 * Host-side work only (schedule build, plan compilation, relabelling);
 * device buffers are created on the first gat_count. */
int gat_plan_create(gat_graph *g, const char *const *adj_strs, int32_t n_patterns,
                    int32_t mode, gat_plan **out);
/* gat_plan_create with options (NULL = defaults). GAT_SCHEDULE_GRAPH picks
 * schedules with this graph's statistics: the reference's Schedule hard-codes
 * its cost-model inputs (schedule.hpp:909-910). */
int gat_plan_create_ex(gat_graph *g, const char *const *adj_strs, int32_t n_patterns,
                       int32_t mode, const gat_plan_options *opt, gat_plan **out);
/* Number of patterns and the schedule record of pattern i. */
int gat_plan_schedule(const gat_plan *p, int32_t i, gat_schedule_data *out, size_t cap);
/* Replaces Executor::perform_search (engine.cuh:232-265) for every pattern
 * of the plan in one persistent launch. counts_out has n_patterns entries.
 * The timed span matches target/pm.cu:60-68 (the search only). */
int gat_count(gat_plan *p, const gat_kernel_config *cfg, uint32_t flags,
              uint64_t *counts_out, gat_count_stats *stats);
/* Capture the count sequence of `cfg` into a hipGraph once, then replay it
 * `replays` times; counts and stats are those of the last replay. */
int gat_count_graph(gat_plan *p, const gat_kernel_config *cfg, int32_t replays,
                    uint64_t *counts_out, gat_count_stats *stats);

/* How pattern i of a plan is counted. In unique and labeled modes the raw
 * count is the number of maps psi from the Schedule-ordered pattern
 * (adj_rows of gat_plan_schedule) into the graph the plan runs on that are
 * injective, send pattern edges to graph edges and satisfy psi(b) < psi(a)
 * for every restriction (a, b) listed here; raw = divisor * count, and
 * gat_count reports raw / divisor * multiplier. With relabeled = 1 that graph
 * is the degree-relabelled one: vertex i is the i-th of the input's vertices
 * ordered by degree, highest first, ties by id. Reference mode reports the
 * reference engine's value, raw itself (divisor and multiplier 1): its
 * restrictions are the reference's filtered ones, and with them the raw
 * count is not the map count above (include/core/schedule.hpp:828-831). No
 * counterpart in the reference, where Schedule::print_schedule
 * (schedule.hpp:931-941) only prints the restrictions. */
typedef struct gat_pattern_count_info {
    uint64_t divisor;
    uint64_t multiplier;
    int32_t relabeled;
    int32_t n_restrictions;
    int32_t restr_first[GAT_MAX_VERTEXES * GAT_MAX_VERTEXES];
    int32_t restr_second[GAT_MAX_VERTEXES * GAT_MAX_VERTEXES];
} gat_pattern_count_info;
int gat_plan_count_info(const gat_plan *p, int32_t i, gat_pattern_count_info *out, size_t cap);

/* One shard of a count: only the embeddings whose Schedule vertex 0 maps to
 * a root in [root_begin, root_end), in vertex ids of the graph the plan runs
 * on (for cliques in unique and labeled modes the root is the clique's
 * highest id). raw_out[i] receives pattern i's raw sum
 * (gat_plan_count_info); the raw sums of shards that partition [0, v_cnt)
 * add up to what gat_count divides by the divisor. No counterpart in the
 * reference, which counts a whole graph per call (engine.cuh:232-265): a
 * shard is the share of one GPU when a count is split across GPUs, and
 * bounds the work of a count on a large graph to a subset of its roots. */
int gat_count_shard(gat_plan *p, const gat_kernel_config *cfg, uint32_t flags,
                    int32_t root_begin, int32_t root_end, uint64_t *raw_out,
                    gat_count_stats *stats);
/* A finer shard for the array store: only the embeddings whose first two
 * Schedule vertices map to a directed edge (root, neighbour) with index in
 * [edge_begin, edge_end) of the CSR the plan runs on (for relabelled plans
 * the degree-ordered CSR: offsets are the prefix sums of the degrees sorted
 * in descending order). Raw sums of a partition of [0, e_directed) add up to
 * divisor * count, like those of gat_count_shard; one heavy root's work can
 * so be split across launches or GPUs. Clique plans on the bitmap store
 * fail with GAT_EUNSUPPORTED (they shard by top vertex). No counterpart in
 * the reference (engine.cuh:246-261 splits the first vertex's range only). */
int gat_count_shard_edges(gat_plan *p, const gat_kernel_config *cfg, uint32_t flags,
                          int64_t edge_begin, int64_t edge_end, uint64_t *raw_out,
                          gat_count_stats *stats);
void gat_plan_free(gat_plan *p);

/* Allocation-free clique count: #K3 .. #K_kmax (kmax in [3, 6]) of a graph
 * already in device memory, written to counts_out[0 .. kmax-3] (host). It
 * runs the kernels of gat_count's bitmap clique store (the vertex of each
 * clique with the highest id is its root; ids in degree-descending order
 * keep the per-vertex lists short) and calls no allocating routine: its
 * counters, per-vertex split array and workgroup-tier lists sit in a static
 * 32 MiB device array, one per GPU, held until the process ends.
 *   d_offsets[v_cnt+1], d_adj: device CSR, rows sorted, symmetric, no
 *   self-loops or duplicates, ids in [0, v_cnt).
 *   stream: a hipStream_t (NULL = the null stream); the call enqueues on it
 *   and waits for it.
 * Limits (GAT_ELIMIT beyond them): v_cnt <= 4,194,304; at most 1,024
 * neighbours below any vertex (no HBM tier); at most 2^20 vertices with more
 * than 512 of them. Calls that overlap on one GPU share its scratch array
 * and race. No counterpart in the reference (engine.cuh:157-216 allocates
 * its storages per Executor): it is for callers that cannot allocate
 * device memory, such as code under HIP graph capture or a serving loop. */
int gat_count_cliques_noalloc(const int64_t *d_offsets, const int32_t *d_adj, int32_t v_cnt,
                              int32_t kmax, const gat_kernel_config *cfg, void *stream,
                              uint64_t *counts_out);

/* Records of the last gat_count run with GAT_FLAG_TIMELINE on this plan, 8
 * int64 each: launch (clique store: 0 main, 1-3 workgroup tiers; array
 * store: 0 main, 1-3 subtask levels, 4 HBM launch, 5 its subtasks), unit
 * (wave of the launch, or workgroup of a clique tier), start and end in
 * wall-clock ticks from the count's first start, shader cycles spent
 * building bit rows and counting pairs (clique store), items taken from the
 * launch's queue and pairs counted (clique store). Writes up to `cap`
 * records, the number available to *n_out and the wall-clock ticks per
 * millisecond to *ticks_per_ms. Load-imbalance evidence; no counterpart in
 * the reference. */
int gat_plan_timeline(const gat_plan *p, int64_t *out, int64_t cap, int64_t *n_out, double *ticks_per_ms);

/* Text dump of a plan's compiled device program (debugging aid). Returns
 * the full length of the text (which may exceed cap) or an error code. */
int gat_plan_describe(const gat_plan *p, char *buf, size_t cap);

/* Default configuration the engine uses for zero fields. */
int gat_default_config(gat_kernel_config *out);

/* ---- learned cost model (tuner) ------------------------------------- */
/* A 3-layer MLP in_dim -> hidden -> hidden -> 1 (ReLU) whose products run
 * on BF16 matrix cores with FP32 accumulation; FP32 master weights, Adam.
 * Replaces the XGBoost regressor of tuning/tuning/role/modeling.py:16-162
 * (fit on measured runtimes, predict to rank candidate configs).
 * Parameters are laid out W1[in][h], b1[h], W2[h][h], b2[h], W3[h], b3. */
typedef struct gat_mlp gat_mlp;
int gat_mlp_create(int32_t in_dim, int32_t hidden, uint64_t seed, gat_mlp **out);
int gat_mlp_param_count(const gat_mlp *m, int64_t *out);
int gat_mlp_get_params(gat_mlp *m, float *out, int64_t count);
int gat_mlp_set_params(gat_mlp *m, const float *in, int64_t count);
/* x: batch x in_dim row-major host array; y_out: batch predictions. */
int gat_mlp_forward(gat_mlp *m, const float *x, int32_t batch, float *y_out);
/* Gradients of the mean squared error against targets t (no update). */
int gat_mlp_gradients(gat_mlp *m, const float *x, const float *t, int32_t batch, float *grads_out,
                      float *loss_out);
/* `steps` full-batch Adam steps; loss_out gets the last step's loss. */
int gat_mlp_train(gat_mlp *m, const float *x, const float *t, int32_t batch, int32_t steps, float lr,
                  float *loss_out);
void gat_mlp_free(gat_mlp *m);

/* ---- device probes --------------------------------------------------- */
/* Number of visible HIP devices (0 when none); never initialises more than
 * hipGetDeviceCount does. */
int gat_device_count(int32_t *out);
/* Plain device-to-device copy of `bytes` through a HIP kernel, timed with
 * hipEvents; used to scale the HBM traffic counters. */
int gat_copy_probe(uint64_t bytes, int32_t reps, double *gbps_out);

#ifdef __cplusplus
}
#endif
#endif /* GAT_H */
