/*
 * cmi.h -- C-ABI of the MI355X interval branch-and-bound engine.
 *
 * Plain C: opaque handles, plain pointers and sizes, status codes instead of
 * exceptions. Every entry point is the flat form of one reference interface
 * on the hot path; each declaration cites the reference file:line it replaces
 * (paths relative to the reference repository root).
 *
 * Ownership: handles are created by *_create and released by the matching
 * *_destroy. View structs (cmi_bound_view, cmi_candidate_view) point into
 * storage owned by the role and stay valid until the next call on that role,
 * as the reference's result spans do (include/cuminlp/backend/backend.hpp:49-50).
 * Threading: a handle is used by one thread at a time; distinct handles on the
 * same context may be used from different threads only if serialised by the
 * caller (all device work of a context is ordered on its one stream).
 * Errors: a non-CMI_OK return leaves the reason in cmi_last_error() (per thread).
 */
#ifndef CMI_H
#define CMI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* == cu::interval<double> layout (16 B), reference kernels.cuh:426-427 */
typedef struct
{
  double lb;
  double ub;
} cmi_ivl;

/* Status codes: one per reference exception class (errors.hpp:11-57,
 * backend.hpp:236-287, cuda_utils.cuh:22-35). */
typedef enum
{
  CMI_OK = 0,
  CMI_SHAPE_MISMATCH = 1,   /* cuminlp::ShapeMismatch */
  CMI_INVALID_CONFIG = 2,   /* cuminlp::InvalidConfiguration */
  CMI_OVER_BUDGET = 3,      /* backend::OverBudgetError / ResourceExhausted */
  CMI_DEVICE_ERROR = 4,     /* CUDAError -> a HIP runtime failure */
  CMI_INVALID_HANDLE = 5,   /* null or destroyed handle (moved-from role) */
  CMI_INVALID_DAG = 6,      /* cuminlp::InvalidDAG */
  CMI_INVALID_PROBLEM = 7,  /* cuminlp::InvalidProblem */
  CMI_ERROR = 8,            /* any other failure */
  CMI_PARSE_ERROR = 9       /* gams::ParseError */
} cmi_status;

/* model::Op, reference model/dag.hpp:12-33 (same order) */
enum
{
  CMI_OP_VAR = 0,
  CMI_OP_CONST,
  CMI_OP_ADD,
  CMI_OP_SUB,
  CMI_OP_MUL,
  CMI_OP_DIV,
  CMI_OP_SQR,
  CMI_OP_NEG,
  CMI_OP_EXP,
  CMI_OP_LOG,
  CMI_OP_SQRT,
  CMI_OP_SIN,
  CMI_OP_COS,
  CMI_OP_TANH,
  CMI_OP_POWN,
  CMI_OP_POW,
  CMI_OP_ABS,
  CMI_OP_MIN,
  CMI_OP_MAX
};

/* model::VarKind and model::Cmp, reference model/problem.hpp:15-26 */
enum
{
  CMI_CONTINUOUS = 0,
  CMI_INTEGER = 1,
  CMI_BINARY = 2
};
enum
{
  CMI_LE = 0,
  CMI_EQ = 1
};

/* region::SlotKind, reference region/slot.hpp:12-18 */
enum
{
  CMI_SLOT_CONTINUOUS = 0,
  CMI_SLOT_INTEGER_PARTITION = 1,
  CMI_SLOT_INTEGER_ENUMERATE = 2,
  CMI_SLOT_BINARY_ENUMERATE = 3
};

/* the three backend roles, reference backend/backend.hpp:81-116 */
enum
{
  CMI_ROLE_SAMPLER = 0,
  CMI_ROLE_BOUNDER = 1,
  CMI_ROLE_ENUMERATOR = 2
};

/* search::StopReason, reference search/budget.hpp:21-30 */
enum
{
  CMI_STOP_CONVERGED = 0,
  CMI_STOP_EXHAUSTED,
  CMI_STOP_ITERATION_LIMIT,
  CMI_STOP_HOST_MEMORY,
  CMI_STOP_ALLOCATION_FAILURE,
  CMI_STOP_DEVICE_MEMORY
};

typedef struct cmi_problem cmi_problem;
typedef struct cmi_ctx cmi_ctx;
typedef struct cmi_tape cmi_tape;
typedef struct cmi_role cmi_role;
typedef struct cmi_server cmi_server;

/* A Problem as flat arrays: the reference's Problem<double> (model/problem.hpp:
 * 195-358) with its ExprDAG (model/dag.hpp:58-93) spelled node by node. in0/in1
 * are read only for ops of arity >= 1 / 2; constant/var_index/int_exp only for
 * Const/Var/PowN nodes (any of the three may be NULL if no node needs it). */
typedef struct
{
  size_t n_nodes;
  const uint8_t* op;
  const uint64_t* in0;
  const uint64_t* in1;
  const double* constant;
  const uint64_t* var_index;
  const int32_t* int_exp;
  size_t n_vars;
  const cmi_ivl* box;
  const uint8_t* var_kind;
  uint64_t objective_root;
  size_t n_constraints;
  const uint64_t* con_root;
  const uint8_t* con_cmp;
  const double* con_rhs;
} cmi_problem_desc;

/* Seeded synthetic QCQP generator (no reference counterpart; it builds the
 * benchmark instances). See qcqp.hpp for the construction. */
typedef struct
{
  size_t n, m, n_int;
  double cont_lb, cont_ub, int_lb, int_ub;
  double pair_density, term_prob, lin_prob;
  int symmetric_pairs;
  double rhs_lo, rhs_hi;
  size_t target_nodes; /* > 0: choose term_prob so the reachable node count is closest to this */
} cmi_qcqp_spec;

typedef struct
{
  size_t n_nodes;
  size_t n_vars;
  size_t n_constraints;
  size_t reachable_nodes; /* reference backend/graph/cost.hpp:38-68 buffer_node_count */
} cmi_problem_info;

typedef struct
{
  size_t n_instr;
  size_t n_slots;
  size_t n_vars;
  size_t n_outputs;
  size_t reachable_nodes;
  uint64_t flops_per_box; /* algorithmic FP64 ops per box of the interval tape */
  size_t bytes_per_box;   /* algorithmic HBM bytes per explicit box: 16 n_vars + 16 + 1 */
  size_t box_groups;      /* waves per explicit-box workgroup (function groups) */
  size_t box_instr;       /* tape words over all groups of the explicit-box code */
  size_t box_slots;       /* LDS slots per lane of the explicit-box code */
  int box_engine;         /* cmi_box_engine in effect */
  int box_compiled;       /* 1 when a compiled explicit-box kernel is loaded */
  double compile_seconds; /* time spent compiling that kernel (0 if none) */
  int compiled_persistent;  /* compiled kernel form: 1 persistent, 0 direct */
  size_t compiled_wgs_per_cu;
  size_t compiled_waves;    /* waves per workgroup of the compiled kernel */
  int extended;             /* the tape uses the extended op set (Div, PowN, Sqrt, Exp, Log, Sin, Cos, Tanh,
                               Abs, Min, Max; Pow as exp(y log x)) */
  int compiled_schedule;    /* cmi_compiled_schedule of the compiled kernel */
  size_t compiled_load_batch; /* variables per load batch of the compiled direct form (0: all at once) */
} cmi_tape_info;

/* Kernel for explicit boxes: the tape interpreter, or the tape compiled to a
 * straight-line gfx950 kernel at run time; AUTO compiles once a call has at
 * least 32,768 boxes (tapes of at most 20,000 words). Results are identical. */
typedef enum
{
  CMI_BOX_AUTO = 0,
  CMI_BOX_INTERPRETED = 1,
  CMI_BOX_COMPILED = 2
} cmi_box_engine;

/* Affine device cost per role, reference backend/cost_model.hpp:26-76 */
typedef struct
{
  size_t sampler_fixed, sampler_per_region;
  size_t bounder_fixed, bounder_per_region;
  size_t enumerator_fixed, enumerator_per_region;
  size_t max_regions; /* most regions one role launch indexes at this sample count */
} cmi_cost_model;

/* OverBudget facts, member for member the reference's OverBudget
 * (backend/backend.hpp:236-262); filled by a build that does not fit, before
 * anything is allocated. */
typedef struct
{
  char role[32];                  /* e.g. "interval bounder" */
  size_t needed_bytes;
  size_t budget_bytes;
  size_t n_regions;
  size_t elements_per_region;     /* > 1 only for the sampler (its samples per region) */
  size_t bytes_per_element;
  char element_breakdown[96];     /* what one element's bytes pay for */
  char element_inventory[64];     /* what the per-element cost grows with */
  cmi_cost_model cost;            /* the problem's affine role costs */
  uint8_t composition_kinds[64];  /* cmi_slot_kind of each slot */
  size_t composition_size;        /* slots in composition_kinds */
  char composition[68];           /* the same kinds spelled, e.g. "CCPE" */
  int has_fan_out;                /* fan_out below is set */
  size_t fan_out_partition_num;
  size_t fan_out_enumerate_cap;
  size_t solve_samples_per_region; /* the solve-wide sample count */
  size_t sampler_bytes;           /* device bytes already held by the sampler */
} cmi_over_budget;

/* BoundResult, reference backend/backend.hpp:68-74 */
typedef struct
{
  const uint8_t* feasible;
  const double* obj_lb;
  size_t n_regions;
} cmi_bound_view;

/* CandidateResult, reference backend/backend.hpp:57-64 */
typedef struct
{
  int found;
  double value;
  const double* point; /* n_vars values; all NaN when !found */
  int64_t index;
} cmi_candidate_view;

/* SearchDriver arguments, reference search/driver.hpp:90-117 plus the
 * GreedyEnumerate policy's FanOutSpec/SearchCalibration (policy/greedy_enum.hpp:
 * 38-44, region/fan_out.hpp:22-57). */
typedef struct
{
  size_t partition_num;   /* FanOutSpec partition width (>= 2) */
  size_t enumerate_cap;   /* FanOutSpec enumerate cap (>= 1) */
  size_t max_cycle_size;  /* slots per iteration, 0 = 64 */
  size_t sample_points;   /* points per region for the sampler */
  uint32_t iter_limit;
  double tolerance;
  size_t budget_bytes;      /* device bytes per role build, 0 = free memory */
  size_t host_budget_bytes; /* 0 = half of MemAvailable */
  int compact_frontier;     /* FrontierPolicy::Compact instead of StopAtBudget */
  int has_known_primal;
  double known_primal;
  /* Composition policy (reference policy/policy.hpp PolicyKind): GreedyEnumerate
   * or WidthFirst over partition_num/enumerate_cap (policy/greedy_enum.hpp,
   * policy/width_first.hpp), or BisectionBudget with 2^bisection_budget regions
   * per node (policy/bisection_budget.hpp), for which enumerate_cap is the
   * largest integer domain enumerated (0 = no ceiling) and partition_num is
   * not read. */
  int policy;               /* cmi_policy */
  size_t bisection_budget;  /* B in [1, 62] for the bisection-budget policies */
} cmi_solve_options;

typedef enum
{
  CMI_POLICY_GREEDY_ENUMERATE = 0,
  CMI_POLICY_BISECTION_BUDGET = 1,
  CMI_POLICY_RELATIVE_BISECTION_BUDGET = 2,
  CMI_POLICY_WIDTH_FIRST = 3
} cmi_policy;

/* SolveOutcome + RunTelemetry, reference search/driver.hpp:36-54 */
typedef struct
{
  double lower_bound;
  double upper_bound;
  int stop_reason;
  int proven_optimal;
  int infeasible;
  int found_incumbent;
  uint64_t iterations;
  uint64_t fathomed;
  uint64_t bounded;
  uint64_t sampled;
  uint64_t enqueued;
  uint64_t pruned_infeasible;
  uint64_t pruned_dominated;
  uint64_t pending;
  double wall_seconds;
  int has_over_budget;         /* stop_reason is CMI_STOP_DEVICE_MEMORY because a role build did not fit */
  cmi_over_budget over_budget; /* its facts (reference driver.hpp:408-416) */
  /* RunTelemetry beyond the counters above (reference report/telemetry.hpp) */
  uint64_t fathomed_points;
  uint64_t witness_rejected;
  uint64_t pruned_duplicate;
  uint64_t viable;             /* pending nodes with lb <= upper_bound at the end */
  uint64_t dropped_viable;     /* compaction evictions the incumbent did not dominate */
  uint64_t dropped_dominated;  /* compaction evictions it did */
  double dropped_floor;        /* least lb among the viable evictions; +inf if none */
  uint64_t history_freed;
  uint64_t history_peak_live;
} cmi_solve_outcome;

const char* cmi_last_error(void);
const char* cmi_version(void);

/* -- problems (host) --------------------------------------------------------- */

/* Build and validate a Problem; replaces Problem construction + validate()
 * (model/problem.hpp:209-351). */
cmi_status cmi_problem_create(const cmi_problem_desc* desc, cmi_problem** out);
cmi_status cmi_problem_qcqp(const cmi_qcqp_spec* spec, uint64_t seed, cmi_problem** out);
cmi_status cmi_problem_info_get(const cmi_problem* p, cmi_problem_info* out);
/* Copy the problem back out as flat arrays; every array argument may be NULL
 * and must otherwise hold n_nodes / n_vars / n_constraints entries. */
cmi_status cmi_problem_export(const cmi_problem* p, uint8_t* op, uint64_t* in0, uint64_t* in1, double* constant,
                              uint64_t* var_index, int32_t* int_exp, cmi_ivl* box, uint8_t* var_kind,
                              uint64_t* objective_root, uint64_t* con_root, uint8_t* con_cmp, double* con_rhs);
void cmi_problem_destroy(cmi_problem* p);

/* -- device context and tapes -------------------------------------------------- */

cmi_status cmi_ctx_create(int device, cmi_ctx** out);
void cmi_ctx_destroy(cmi_ctx* ctx);

/* Compile the problem's DAG to a tape and upload it; replaces GraphBuilder's
 * DAG walk (backend/graph/builder.cuh:261-514). */
cmi_status cmi_tape_create(cmi_ctx* ctx, const cmi_problem* p, cmi_tape** out);
cmi_status cmi_tape_info_get(const cmi_tape* t, cmi_tape_info* out);
void cmi_tape_destroy(cmi_tape* t);
/* Split the explicit-box code of tape `t` (compiled from `p`) into k function
 * groups, one wave each over the same 64 boxes (0 = chosen from the tape's
 * size). k is clamped to [1, min(16, 1 + n_constraints)] and lowered until the
 * workgroup fits in LDS. Results do not depend on k. No reference
 * counterpart: the reference launches one kernel per DAG node
 * (backend/graph/builder.cuh:261-514); this is its wave-shape knob. Must not
 * overlap a launch on the tape. */
cmi_status cmi_tape_set_box_groups(cmi_tape* t, const cmi_problem* p, size_t k);
/* Select the explicit-box kernel (cmi_box_engine). CMI_BOX_COMPILED compiles
 * now; a compiler failure returns CMI_DEVICE_ERROR with the compiler log in
 * cmi_last_error(). No reference counterpart (the reference replays a CUDA
 * graph of per-node kernels, backend/graph/replay.cuh:493); must not overlap a
 * launch on the tape. */
cmi_status cmi_tape_set_box_engine(cmi_tape* t, int engine);
/* Shape of the compiled explicit-box kernel. persistent = 1 walks the 64-box
 * tiles with workgroups that stay resident (wgs_per_cu per CU) and stage the
 * next tile into LDS while evaluating the current one (used only when two
 * tiles of the problem's variables fit in LDS); 0 launches one workgroup per
 * tile. waves = waves per workgroup (0 = one per function group, else at most
 * 8): wave v runs the function groups v, v + waves, ..., interleaving their
 * statements. Results do not depend on the shape. */
cmi_status cmi_tape_set_compiled_shape(cmi_tape* t, int persistent, size_t waves, size_t wgs_per_cu);
/* Instruction scheduler the run-time compiler orders the compiled kernel's
 * code with: the default, occupancy-first one, or the one that maximises
 * instruction-level parallelism. The kernel source, and so every result, is
 * the same either way. */
typedef enum
{
  CMI_SCHED_DEFAULT = 0,
  CMI_SCHED_MAX_ILP = 1
} cmi_compiled_schedule;
/* Code generation of the compiled kernel: its scheduler (cmi_compiled_schedule)
 * and, for the direct form, its load batching: load_batch = 0 issues every
 * variable load at the start of the wave; k in [1, 64] issues them k
 * variables at a time in the order the statements first read them, each
 * batch as the statements start on the one before it. Recompiles when the
 * compiled engine is in effect. Results do not depend on either. A call that
 * asks for root intervals runs a second compiled kernel that writes them,
 * compiled on the first such call; without them the kernel computes only
 * what its outputs read. No reference counterpart; must not overlap a launch
 * on the tape. */
cmi_status cmi_tape_set_compiled_codegen(cmi_tape* t, int schedule, size_t load_batch);
/* One explicit-box kernel configuration and its median kernel time. */
typedef struct
{
  int engine;        /* cmi_box_engine: INTERPRETED or COMPILED */
  size_t groups;     /* function groups */
  size_t waves;      /* waves per workgroup (compiled) */
  int persistent;    /* compiled form */
  size_t wgs_per_cu; /* persistent form: workgroups per CU */
  int schedule;      /* compiled: cmi_compiled_schedule */
  size_t load_batch; /* compiled direct form: variables per load batch (0: all at once) */
  double kernel_ms;
} cmi_box_candidate;

/* Autotuner: time the interpreter and the compiled shapes (function groups,
 * waves per workgroup, persistent or direct, workgroups per CU, and the direct
 * form's scheduler and load batching) on n_boxes synthetic boxes inside the
 * bounds of `p` (the problem the tape was built from), streamed from HBM
 * (rotating copies of up to 640 MB in all) in sustained runs of 64
 * back-to-back launches; time the four fastest again, leave the tape on the
 * fastest of those, and report up to `cap` candidates in `results`
 * (*n_results = how many ran; kernel_ms per launch). Takes seconds (each
 * compiled shape is compiled once). No reference counterpart: the reference
 * has one fixed launch shape per node kernel (backend/graph/kernels.cuh:177-281). */
cmi_status cmi_tape_autotune(cmi_tape* t, const cmi_problem* p, size_t n_boxes, cmi_box_candidate* results, size_t cap,
                             size_t* n_results);

/* One compiled explicit-box kernel, for the host-only views below. */
typedef struct
{
  size_t groups;     /* function groups as cmi_tape_set_box_groups makes them; 0 = those of a new tape */
  int persistent;    /* 1 persistent form (INVALID_CONFIG if its tiles do not fit in LDS), 0 direct, -1 a new
                        tape's form */
  size_t waves;      /* as for cmi_tape_set_compiled_shape */
  int schedule;      /* cmi_compiled_schedule */
  size_t load_batch; /* as for cmi_tape_set_compiled_codegen */
  int roots;         /* 1: the kernel that writes root intervals */
} cmi_compiled_shape;
/* Host-only: the HIP source of the compiled explicit-box kernel of `p`'s tape
 * in `shape`, NUL-terminated into `buf` (up to `cap` bytes); *len receives
 * the full length without the NUL. */
cmi_status cmi_tape_compiled_source(const cmi_problem* p, const cmi_compiled_shape* shape, char* buf, size_t cap,
                                    size_t* len);
/* Host-only: compile that source with the run-time compiler, exactly as
 * CMI_BOX_COMPILED does, and copy the gfx950 code object into `buf` (when
 * non-NULL and cap >= *len; SHAPE_MISMATCH otherwise); *len receives its size. */
cmi_status cmi_tape_compiled_object(const cmi_problem* p, const cmi_compiled_shape* shape, void* buf, size_t cap,
                                    size_t* len);
/* Compile the tape on the host only (no device) and report its shape; also
 * writes the listing into `listing` (up to `cap` bytes, NUL-terminated) when
 * non-NULL. Same checks and errors as cmi_tape_create. */
cmi_status cmi_tape_compile_info(const cmi_problem* p, cmi_tape_info* out, char* listing, size_t cap);
/* Host-only compile into k function groups (k <= 1: one group, the tape the
 * decode-mode roles run; k is clamped to min(16, 1 + n_constraints)). Writes
 * the shape into `out` (box_groups = the k used), the 16-byte tape words into
 * `words` when non-NULL (SHAPE_MISMATCH if cap_words < out->box_instr) and
 * the box_groups + 1 group offsets into `group_begin` when non-NULL. Lets a
 * host-side interpreter replay exactly the words a device runs. No reference
 * counterpart (the reference has no tape). */
cmi_status cmi_tape_compile_words(const cmi_problem* p, size_t k, void* words, size_t cap_words,
                                  uint32_t* group_begin, cmi_tape_info* out);

/* Evaluate explicit boxes (box e, variable v: [lb[v*ld+e], ub[v*ld+e]]):
 * feasible[e] and obj_lb[e] as the bounder defines them (kernels.cuh:325-372),
 * obj_ub[e] and roots[k*n_boxes+e] (k = 0 objective, 1.. constraints) when
 * non-NULL. Host pointers; synchronous. */
cmi_status cmi_eval_boxes(cmi_tape* t, const double* lb, const double* ub, size_t ld, size_t n_boxes,
                          uint8_t* feasible, double* obj_lb, double* obj_ub, cmi_ivl* roots);
/* Same with device pointers, enqueued on `stream` (a hipStream_t, NULL = the
 * context's stream); returns without synchronising. */
cmi_status cmi_eval_boxes_device(cmi_tape* t, const double* lb, const double* ub, size_t ld, size_t n_boxes,
                                 uint8_t* feasible, double* obj_lb, double* obj_ub, cmi_ivl* roots, void* stream);
/* Round-to-nearest point evaluation (the point graph's op path, ops.cuh
 * apply(T,T) overloads): x[v*ld+e]; host pointers. */
cmi_status cmi_eval_points(cmi_tape* t, const double* x, size_t ld, size_t n_points, uint8_t* feasible, double* obj,
                           double* roots);

/* -- backend roles ------------------------------------------------------------ */

/* RegionBackendFactory::cost_model, backend/backend.hpp:200-201 */
cmi_status cmi_cost_model_get(const cmi_tape* t, size_t samples_per_region, cmi_cost_model* out);
/* RegionBackendFactory::build_subdivision for one role, backend/backend.hpp:213-218.
 * On CMI_OVER_BUDGET `facts` (if non-NULL) is filled and nothing was allocated. */
cmi_status cmi_role_build(cmi_tape* t, int role_kind, const uint8_t* slot_kinds, size_t slot_count, size_t n_regions,
                          size_t samples_per_region, size_t budget_bytes, cmi_role** out, cmi_over_budget* facts);
/* RegionBounder::bound, backend/backend.hpp:99 */
cmi_status cmi_bound(cmi_role* r, const cmi_ivl* box, size_t n_vars, const size_t* var_ids, const size_t* widths,
                     size_t slot_count, cmi_bound_view* out);
/* RegionSampler::sample, backend/backend.hpp:87 */
cmi_status cmi_sample(cmi_role* r, const cmi_ivl* box, size_t n_vars, const size_t* var_ids, const size_t* widths,
                      size_t slot_count, uint64_t salt, cmi_candidate_view* out);
/* RegionEnumerator::enumerate, backend/backend.hpp:114 */
cmi_status cmi_enumerate(cmi_role* r, const cmi_ivl* box, size_t n_vars, const size_t* var_ids,
                         const size_t* widths, size_t slot_count, cmi_candidate_view* out);
size_t cmi_role_size(const cmi_role* r); /* n_regions / n_samples / n_points */

/* The sampler's search for the best sampled point of a region, without a role
 * and without allocating: no hipMalloc/hipFree and no allocating library
 * call. No counterpart in the reference; it is for callers that cannot
 * allocate device memory (code under graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Arguments and results are cmi_sample's
 * on a sampler built for (slot_kinds, prod(widths) regions,
 * samples_per_region), with the candidate written to caller-owned host memory
 * (found, value, index; point holds n_vars doubles). Staging, per-tile
 * partial minima and the done counter live in a 4 MiB file-scope device array
 * per GPU, held until the process ends and kept between calls (the call
 * clears what must start at zero). A call whose scratch would exceed it (about
 * 200,000 tiles of 64 points, fewer with many variables) returns
 * CMI_INVALID_CONFIG before anything runs. Calls on one device are
 * serialised inside the process; calls that overlapped on one GPU would race
 * on the shared array. */
cmi_status cmi_sample_noalloc(cmi_tape* t, const uint8_t* slot_kinds, size_t slot_count, const cmi_ivl* box,
                              size_t n_vars, const size_t* var_ids, const size_t* widths, size_t samples_per_region,
                              uint64_t salt, int* found, double* value, int64_t* index, double* point);
void cmi_role_destroy(cmi_role* r);

/* -- aggregate (k-child) bounder ------------------------------------------------ */

/* AggregateBound, reference aggregate/bound.hpp:42-64: min over a child's
 * unexcluded subregions of the objective's lower bound, max of the upper
 * bound (hull_ub, never an incumbent), and whether any subregion survived.
 * An infeasible child is {+inf, +inf, 0}. */
typedef struct
{
  double lb;
  double hull_ub;
  int feasible;
} cmi_aggregate_bound;

/* AggregateBackendFactory::build_bounder, reference
 * backend/aggregate/factory.cuh and replay.cuh:49-118: a bounder for launch
 * assignments with these slot kinds, N = n_regions subregions per region in k
 * children of M = N / k (k must divide N), taking up to max_batch regions per
 * call. Its device bytes do not grow with N beyond one u32 per 64 subregions.
 * CMI_OVER_BUDGET with `facts` before allocating. The result is a cmi_role
 * that only cmi_bound_children accepts. */
cmi_status cmi_aggregate_build(cmi_tape* t, const uint8_t* slot_kinds, size_t slot_count, size_t n_regions, size_t k,
                               size_t max_batch, size_t budget_bytes, cmi_role** out, cmi_over_budget* facts);
/* AggregateBounder::bound_children, reference aggregate/bound.hpp:88-91, for
 * n_batch regions in one launch (the reference builds one region per launch,
 * backend/aggregate/replay.cuh:255-270): region i is boxes[i*n_vars ..] with
 * launch slots var_ids[i*slot_count ..], widths[i*slot_count ..] (product N;
 * the branch slots last, so child c owns subregions [c*M, (c+1)*M)). Writes
 * n_batch * k bounds to `out`, region-major. */
cmi_status cmi_bound_children(cmi_role* r, size_t n_batch, const cmi_ivl* boxes, size_t n_vars, const size_t* var_ids,
                              const size_t* widths, size_t slot_count, cmi_aggregate_bound* out);

/* -- the k-child (aggregate) search --------------------------------------------- */

/* SelectionRule, reference aggregate/selection.hpp:15-125: which pending
 * child is explored next. Every rule orders by lb until an incumbent exists. */
typedef enum
{
  CMI_SELECT_LEAST_LOWER_BOUND = 0,
  CMI_SELECT_REJECT_INDEX = 1,       /* max (f_best - lb) / (hull_ub - lb), re-keyed on every new incumbent */
  CMI_SELECT_REJECT_INDEX_STALE = 2, /* the same key, fixed at enqueue time */
} cmi_selection_rule;

/* The aggregate search's settings (reference source/gams/aggregate_solve.cu:
 * 133-143 flags, 298-330 shape). */
typedef struct
{
  size_t branch_fan_out;    /* k: children per node, a power of two >= 2 (--branch-factor, default 4) */
  size_t partition_budget;  /* N: subregions bounded per node, a power of two > k (--partition-budget) */
  size_t sample_fan_out;    /* subregions sampled per node, a power of two > k */
  size_t sample_points;     /* points drawn per sampled subregion, >= 1 (default 1) */
  uint32_t iter_limit;
  double tolerance;         /* gap closed when gub - bound <= tolerance * max(1, |gub|) */
  int selection;            /* cmi_selection_rule */
  int relative_width;       /* the refine's bisections by width relative to the node's own */
  size_t budget_bytes;      /* device bytes for both roles, 0 = free memory; the sampler may use a tenth */
  size_t host_budget_bytes; /* 0 = half of MemAvailable; on reaching it the search stops (never compacts) */
  int has_known_primal;
  double known_primal;      /* a sound upper bound on the minimum, seeding the incumbent value */
  int kernel;               /* cmi_box_engine of the bound launches: AUTO (interpreted until about a tenth
                               of a second of work, then compiled), INTERPRETED or COMPILED; same results */
} cmi_aggregate_options;

/* Defaults: k 4, N 1024, sample fan-out 8, 1 sample, iter_limit 1000000,
 * tolerance 1e-6, least-lower-bound. */
void cmi_aggregate_options_default(cmi_aggregate_options* o);

/* The reference's device fit of N and the sample fan-out
 * (source/gams/aggregate_solve.cu:298-315, aggregate/cost.hpp:22-99): a
 * device budget of 0.8 * free_device_bytes split 9:1 between bounder and
 * sampler, N the largest power of two whose per-subregion cost
 * (16 * buffer nodes + 33 bytes) fits the bounder's share, the sample
 * fan-out the largest power of two whose (8 * buffer nodes + 25) *
 * sample_points bytes fit the sampler's, at least max(2, 2k). The device
 * search itself holds no per-subregion buffer, so these sizes are the
 * reference's choice, not this engine's memory limit. */
cmi_status cmi_aggregate_fit(const cmi_problem* p, size_t free_device_bytes, size_t branch_fan_out,
                             size_t sample_points, size_t* partition_budget, size_t* sample_fan_out);

/* AggregateDriver::solve, reference aggregate/driver.hpp:98-462: best-first
 * branch and bound in which each node is split k ways on its widest live
 * variable and the device bounds each child over N / k subregions in one
 * launch (the least lower bound and the greatest upper bound over those not
 * proven infeasible), a child's bound clamped to its parent's; the sampler
 * draws sample_points points in each of sample_fan_out subregions of the
 * node. A node with no live variable is a point and is evaluated on the host.
 * `out` counts children in `bounded`; `pending`/`viable` are children.
 * *rule_engaged (optional) is 0 only when a re-keying rule never had a
 * sampled incumbent, so it explored in lb order throughout. */
cmi_status cmi_aggregate_solve(cmi_ctx* ctx, const cmi_problem* p, const cmi_aggregate_options* o,
                               cmi_solve_outcome* out, double* best_point, int* rule_engaged);

/* -- GAMS scalar-format models (host) ----------------------------------------- */

typedef struct cmi_gams_model cmi_gams_model;

/* gams::ParseOptions, reference include/cuminlp/gams.hpp:21-50 */
typedef struct
{
  double default_bound;    /* box half-width for a variable with no finite bound (1e6) */
  int fold_fixed_to_const; /* x.fx variables become literals (1) */
  int reject_discrete;     /* refuse Binary/Integer declarations (0) */
} cmi_gams_options;

/* gams::ErrorKind, reference gams.hpp:59-70 (same order) */
enum
{
  CMI_GAMS_SYNTAX = 0,
  CMI_GAMS_NON_SCALAR,
  CMI_GAMS_DISCRETE,
  CMI_GAMS_UNSUPPORTED_FUNCTION,
  CMI_GAMS_UNREPRESENTABLE,
  CMI_GAMS_IO
};

/* gams::Sense, reference gams.hpp:14-19 */
enum
{
  CMI_SENSE_MIN = 0,
  CMI_SENSE_MAX = 1
};

/* ParseError's line and kind, reference gams.hpp:72-83 */
typedef struct
{
  int line;
  int kind;
} cmi_gams_error;

/* ParsedModel's metadata, reference gams.hpp:85-116 */
typedef struct
{
  int sense;               /* CMI_SENSE_*; under MAX the objective is already negated */
  size_t n_vars;
  size_t n_constraints;
  size_t n_warnings;
  int objvar_kept;         /* the objective variable stayed a search dimension */
  int has_initial_point;   /* the file gave x.l levels */
} cmi_gams_info;

void cmi_gams_options_default(cmi_gams_options* o);
/* gams::parse / gams::parse_file, reference gams.hpp:118-126 and
 * source/gams/frontend.cpp:1592-1681. On CMI_OK, *problem is a new validated
 * problem (release with cmi_problem_destroy) and *model its names, sense,
 * warnings and initial point (release with cmi_gams_destroy). On
 * CMI_PARSE_ERROR nothing is allocated, `err` (may be NULL) receives the line
 * (0 = whole model) and kind, and cmi_last_error() the message. `options` may
 * be NULL for the defaults. */
cmi_status cmi_gams_parse(const char* text, size_t len, const cmi_gams_options* options, cmi_problem** problem,
                          cmi_gams_model** model, cmi_gams_error* err);
cmi_status cmi_gams_parse_file(const char* path, const cmi_gams_options* options, cmi_problem** problem,
                               cmi_gams_model** model, cmi_gams_error* err);
cmi_status cmi_gams_info_get(const cmi_gams_model* m, cmi_gams_info* out);
/* ParsedModel::var_names[i] / constraint_names[i] / warnings[i]; NULL when i
 * is out of range. The strings live until cmi_gams_destroy. */
const char* cmi_gams_var_name(const cmi_gams_model* m, size_t i);
const char* cmi_gams_constraint_name(const cmi_gams_model* m, size_t i);
const char* cmi_gams_warning(const cmi_gams_model* m, size_t i, int* line);
/* ParsedModel::initial_point (n = n_vars values); CMI_INVALID_CONFIG if the
 * file gave no levels. */
cmi_status cmi_gams_initial_point(const cmi_gams_model* m, double* out, size_t n);
void cmi_gams_destroy(cmi_gams_model* m);
/* The solver's RESULT line, reference source/gams/solve.cu:836-867:
 * "RESULT\tsense=min|max\tprimal=<v|none>\tdual=<v|none>" (%.17g) from the
 * minimisation bounds a solve reports (cmi_solve_outcome lower_bound /
 * upper_bound), negated back under CMI_SENSE_MAX. Writes at most cap bytes
 * including the terminating NUL; *len (may be NULL) gets the full length. */
cmi_status cmi_gams_result_line(int sense, double lower, double upper, char* buf, size_t cap, size_t* len);

/* -- memory-fit resolver (reference config/problem_profile.hpp,
 * config/footprint.hpp, config/catalogue.hpp, config/resolve.hpp,
 * config/run_spec.hpp). Host arithmetic only: no device is touched. -------- */

/* ProblemProfile, reference config/problem_profile.hpp:21-32 */
typedef struct
{
  size_t num_binary, num_integer, num_continuous;
  size_t largest_integer_domain; /* integer points in the widest integer box, 0 if none */
  int objvar_kept;               /* a GAMS model whose objective variable stayed a variable */
} cmi_problem_profile;

enum
{
  CMI_SOURCE_AUTO = 0,      /* rules selected by the problem's variable kinds */
  CMI_SOURCE_NAMED = 1,     /* rules named by the caller */
  CMI_SOURCE_OVERRIDDEN = 2 /* at least one number given by the caller */
};

typedef struct
{
  const char* rules;        /* roster row ("all-binary", "discrete", "mixed-binary", "mixed-all-small",
                               "mixed-all-large"), NULL = select by the problem's variable kinds */
  int objvar_kept;          /* cmi_gams_info.objvar_kept of the model, 0 for a hand-built problem */
  size_t free_device_bytes; /* free device memory to fit against, 0 = no budget (the rows' fallbacks) */
  /* overrides, 0 = resolve */
  size_t partition_num, enumerate_cap, sample_points, max_cycle_size;
} cmi_resolve_request;

typedef struct
{
  size_t partition_num, enumerate_cap, sample_points, max_cycle_size;
  size_t budget_bytes;    /* what the fit spent against: 0.67 of free_device_bytes */
  size_t footprint_bytes; /* device bytes of the worst composition the shape can reach */
  int source;             /* CMI_SOURCE_* */
  char rules[32];         /* the roster row used */
} cmi_resolved_shape;

/* profile_problem, reference config/problem_profile.hpp:43-70 (objvar_kept = 0). */
cmi_status cmi_problem_profile_get(const cmi_problem* p, cmi_problem_profile* out);
/* Free and total device memory of the context's device (hipMemGetInfo). */
cmi_status cmi_ctx_memory(cmi_ctx* ctx, size_t* free_bytes, size_t* total_bytes);
/* worst_composition_footprint, reference config/footprint.hpp:61-97: the
 * device bytes (the roles' cost model, cmi_cost_model_get, at sample_points)
 * of the largest composition any box of the search reaches with cap slots.
 * profile may be NULL (then it is taken from p). */
cmi_status cmi_worst_footprint(const cmi_problem* p, const cmi_problem_profile* profile, size_t cap,
                               size_t partition_num, size_t enumerate_cap, size_t sample_points, size_t* bytes);
/* auto_max_cycle_size, reference config/footprint.hpp:99-149: the widest
 * slot cap, at most ceiling, whose worst composition fits budget_bytes; 0 if
 * none does. */
cmi_status cmi_auto_max_cycle_size(const cmi_problem* p, const cmi_problem_profile* profile, size_t partition_num,
                                   size_t enumerate_cap, size_t sample_points, size_t budget_bytes, size_t ceiling,
                                   size_t* out);
/* explain_over_budget, reference config/resolve.hpp:247-330: a role build that
 * did not fit (cmi_solve_outcome.over_budget, or cmi_role_build's facts),
 * explained with the widest slot cap that fits its budget at the solve's
 * fan-out, every role of a composition charged. Writes at most cap bytes
 * including the NUL; *len (may be NULL) gets the full length. */
cmi_status cmi_explain_over_budget(const cmi_problem* p, int objvar_kept, const cmi_over_budget* facts,
                                   size_t partition_num, size_t enumerate_cap, size_t sample_points, char* buf,
                                   size_t cap, size_t* len);
/* select_policy + resolve_shape + resolve, reference config/catalogue.hpp:182-245,
 * config/resolve.hpp:95-218, config/run_spec.hpp:122-166. An unknown or
 * inapplicable named row is CMI_INVALID_CONFIG. */
cmi_status cmi_resolve_shape(const cmi_problem* p, const cmi_resolve_request* req, cmi_resolved_shape* out);

/* -- search ------------------------------------------------------------------- */

void cmi_solve_options_default(cmi_solve_options* o);
/* SearchDriver<double>::solve with GreedyEnumerate over the device roles,
 * search/driver.hpp:137-537. best_point (n_vars) may be NULL. */
cmi_status cmi_solve(cmi_ctx* ctx, const cmi_problem* p, const cmi_solve_options* o, cmi_solve_outcome* out,
                     double* best_point);
/* The same solves for many problems, interleaved: every step runs one search
 * iteration of every unfinished solve in one sampler launch and one bounder
 * launch (the batched form of AggregateBounder::bound_children's span API,
 * aggregate/bound.hpp:85-100). Per-solve results equal cmi_solve's, including
 * a solve whose role builds do not fit budget_bytes (it stops with
 * CMI_STOP_DEVICE_MEMORY and its facts; the others go on). Host budget: a
 * non-zero host_budget_bytes is each solve's own, as in cmi_solve; 0 reads
 * half of MemAvailable once and gives each of the n solves an equal share.
 * With budget_bytes 0 the free device memory is read on the first step and
 * every 32nd step after it. */
cmi_status cmi_solve_batch(cmi_ctx* ctx, size_t n, const cmi_problem* const* problems, const cmi_solve_options* o,
                           cmi_solve_outcome* outs, double* const* best_points);

/* -- serving (no reference counterpart: the reference solves one problem per
 * process; these run a request stream through the same searches) ----------- */

typedef struct
{
  size_t max_active;       /* solves on the device at once (continuous-batching slots), >= 1 */
  size_t hbm_budget_bytes; /* admission budget for the admitted solves' device state, 0 = unlimited */
  int realtime;            /* 1: request i arrives arrival[i] s after the call starts; 0: all at t = 0 */
} cmi_server_options;

typedef struct
{
  uint64_t steps;       /* lock-step iterations run */
  uint64_t launches;    /* kernel launches issued */
  size_t peak_active;   /* most solves resident at once */
  double wall_seconds;  /* first admission to last completion */
  size_t peak_resident_bytes; /* largest sum of the resident solves' admission footprints (0 for serve_plain) */
  double kernel_seconds;  /* device time of the steps' decode-mode launches: HIP events around every 8th
                             step's launches, scaled by FP64 work to all steps (0 for serve_plain) */
  double kernel_fp64_ops; /* their algorithmic FP64 work: interval ops per region, DAG ops per point */
} cmi_serve_stats;

/* cmi_solve_batch, also reporting its steps, launches, wall time and the
 * device time and FP64 work of its decode-mode launches (stats may be NULL;
 * peak_active is the batch size, peak_resident_bytes 0). No reference
 * counterpart: it instruments cmi_solve_batch. */
cmi_status cmi_solve_batch_stats(cmi_ctx* ctx, size_t n, const cmi_problem* const* problems,
                                 const cmi_solve_options* o, cmi_solve_outcome* outs, double* const* best_points,
                                 cmi_serve_stats* stats);

/* Continuous batching: requests are admitted as they arrive while a slot and
 * the HBM budget are free (a solve is charged its tape plus the memory-fit
 * resolver's worst-composition footprint at the options' shape, see
 * cmi_worst_footprint; the first resident solve is always admitted); every step advances every resident solve by one
 * search iteration in one launch per role kind; a finished solve frees its
 * slot at once. finish[i] (seconds since the call started) and latency[i]
 * (finish - arrival) may be NULL. Per-request results equal cmi_solve's, as
 * for cmi_solve_batch. A request is priced (tape + worst-composition role
 * footprint) from a host compile before any device memory is taken; its tape
 * is uploaded only once it is admitted. Host budget: a non-zero
 * host_budget_bytes is each solve's own; 0 reads half of MemAvailable once
 * and gives each of the max_active slots an equal share. Calls that overlap
 * on one context are not supported. */
cmi_status cmi_serve_stream(cmi_ctx* ctx, size_t n, const cmi_problem* const* problems, const double* arrival,
                            const cmi_solve_options* o, const cmi_server_options* so, cmi_solve_outcome* outs,
                            double* finish, double* latency, cmi_serve_stats* stats);
/* The plain baseline of the same stream: one request at a time, in arrival
 * order, each through cmi_solve's one-launch-per-role driver. */
cmi_status cmi_serve_plain(cmi_ctx* ctx, size_t n, const cmi_problem* const* problems, const double* arrival,
                           const cmi_solve_options* o, cmi_solve_outcome* outs, double* finish, double* latency,
                           cmi_serve_stats* stats);

#ifdef __cplusplus
}
#endif

#endif /* CMI_H */
