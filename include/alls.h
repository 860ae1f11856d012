/*
 * alls.h -- C ABI of liballs.so, MI355X (gfx950) A-LLS finite-element
 * stiffness assembly.
 *
 * Inputs are node coordinates, element connectivity and material constants.
 * The output is a CSR triple (row_ptr, col_ind, val).  Its pattern is
 * bit-identical to the reference's host std::set pattern, and its values
 * match the reference's serial element-order assembly.
 *
 * The reference (zhou1533/A-LLS-Matrix-Parallel-Assembly) has no library or
 * FFI.  Each entry point below names the reference code it replaces.  Keys:
 *   Acc = ALLS_Quad2D_Assembly_Accuracy_Analysis.cu
 *   T3  = ALLS_T3_Unstructed_Elements_Assemble_Performance.cu
 *   Mix = ALLS_Q4_T3_Mixed_Elements_Assembly_Performance.cu
 *   Hex = ALLS_Hex20_3D_Assembly_Performance.cu
 *
 * Conventions:
 *  - Errors are status codes; the library never exits (the reference's
 *    CHECK_CUDA_ERROR calls exit(), Acc:15-23).
 *  - No global state.  One context per host thread; all work is ordered on
 *    the context's HIP stream.  The reference keeps its pool pointers in
 *    __device__ globals (Acc:384-385, 790-791), which makes it non-reentrant.
 *  - Output CSR buffers are device memory owned by the library.  Release them
 *    with alls_csr_free.  Passing an alls_csr that already holds large-enough
 *    buffers reuses them.
 */
#ifndef ALLS_H
#define ALLS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    ALLS_OK = 0,
    ALLS_E_INVALID = 1,       /* bad argument or malformed mesh (node id out of range, ...) */
    ALLS_E_OOM = 2,           /* device or host allocation failed */
    ALLS_E_POOL_OVERFLOW = 3, /* a row outgrew the list directory after all retries */
    ALLS_E_HIP = 4,           /* HIP runtime error (message in alls_ctx_last_error) */
    ALLS_E_BUSY = 5,          /* admission refused: request exceeds the HBM budget */
    ALLS_E_NO_DEVICE = 6,     /* no usable GPU */
    ALLS_E_UNSUPPORTED = 7
} alls_status;

/* Element kinds.  Q4=0 and T3=1 are the reference's ElementType values (T3:37-40). */
typedef enum { ALLS_Q4 = 0, ALLS_T3 = 1, ALLS_HEX20 = 2 } alls_elem_type;
typedef enum { ALLS_MEM_HOST = 0, ALLS_MEM_DEVICE = 1 } alls_mem;
/* Precision of Hex20 element integration.  Both forms accumulate and
 * consolidate in FP64.  F32_MFMA contracts the Gauss-point gradients on FP32
 * matrix cores and stores FP32 element contributions. */
typedef enum { ALLS_PREC_F64 = 0, ALLS_PREC_F32_MFMA = 1 } alls_precision;

typedef struct alls_ctx alls_ctx;
typedef struct alls_pattern alls_pattern;
typedef struct alls_coloring alls_coloring;
typedef struct alls_server alls_server;

/* Mesh in structure-of-arrays form.  It replaces the reference's Element AoS,
 * which copies node coordinates into every element (Acc:51-56, T3:53-60,
 * Hex:49-55), and its compile-time material constants (Acc:39-41, Hex:43-44). */
typedef struct {
    int32_t dim;              /* 2 (Q4/T3/mixed) or 3 (Hex20); also DOF per node */
    int32_t nodes_per_elem;   /* stride of conn: 4 (Q4, or mixed Q4/T3), 3 (T3), 20 (Hex20) */
    int64_t n_nodes;
    int64_t n_elems;
    const double*  coords;    /* [n_nodes*dim], interleaved x,y(,z) */
    const int32_t* conn;      /* [n_elems*nodes_per_elem]; unused trailing slots -1 (T3 in a mixed mesh) */
    const uint8_t* elem_type; /* [n_elems] alls_elem_type, or NULL => uniform_type */
    int32_t uniform_type;
    int32_t mem;              /* alls_mem of coords, conn, elem_type and E_per_elem */
    double E;                 /* Young's modulus, used when E_per_elem == NULL */
    const double* E_per_elem; /* [n_elems] or NULL */
    double nu;                /* Poisson ratio */
    double thickness;         /* plane-stress thickness (2D) */
} alls_mesh;

typedef struct {
    int32_t precision;        /* alls_precision (Hex20 only) */
    int32_t chunk_slots;      /* entries per list-pool chunk; 0 => tuned default; at most 16384.
                                 Hex20: a multiple of 20 lets the lists carry one element id per
                                 20-entry strip; other sizes key every entry (same results) */
    int32_t block_threads;    /* threads per block of the insert kernel; 0 => tuned default */
    int32_t row_ptr_i64;      /* 1 => always int64 row_ptr; otherwise int64 only when nnz >= 2^31 */
    int32_t timing;           /* 1 => time each stage with hipEvents (alls_last_stats); 0 => the
                                 pipeline records no events between its stages */
    int32_t grid_per_cu;      /* insert kernel grid: > 0 => persistent, this many blocks per CU looping
                                 over the elements; 0 => one thread (2D) or one wave (Hex20) per element */
    int32_t autotune;         /* 1 => take block_threads / chunk_slots / grid_per_cu from the context's
                                 autotuner (tuned on first use for the mesh's kind and size class) */
    int32_t val_f32;          /* 1 => Hex20 only: the output values are float (summed in FP64, stored
                                 as FP32), as the reference's Hex20 path returns them
                                 (Hex:260, Hex:601-608); alls_assemble and alls_assemble_noalloc only */
} alls_opts;

/* CSR triple in device memory.  Zero-initialise before first use. */
typedef struct {
    int64_t n_rows;
    int64_t nnz;
    int32_t row_ptr_is_i64;   /* row_ptr element type: int32 (0) or int64 (1) */
    int32_t pattern_id;       /* library bookkeeping: cached pattern whose row_ptr/col_ind
                                 these buffers already hold (0 = none) */
    void*    row_ptr;         /* device [n_rows+1] */
    int32_t* col_ind;         /* device [nnz] */
    double*  val;             /* device [nnz]; float [nnz] when val_is_f32 */
    int64_t cap_rows;         /* capacity of row_ptr in rows (library bookkeeping) */
    int64_t cap_nnz;          /* capacity of col_ind/val in entries */
    int32_t val_is_f32;       /* val holds floats (opts.val_f32) */
    int32_t _pad;
} alls_csr;

/* Per-call statistics of the last assembly on a context. */
typedef struct {
    double ms_total;          /* device time of the whole pipeline (hipEvents) */
    double ms_init;           /* list header reset */
    double ms_insert;         /* element integration + list insertion */
    double ms_count;          /* per-row sort + unique count */
    double ms_scan;           /* device prefix scan -> row_ptr */
    double ms_write;          /* per-row sort + merge -> col_ind, val (when the pipeline ran as a
                                 captured graph: the whole graph, which has no stage events) */
    int64_t n_elems;
    int64_t n_rows;
    int64_t nnz;
    int64_t list_entries;     /* entries appended to the row lists */
    int64_t pool_chunks_used;
    int64_t pool_chunks_cap;
    int64_t scratch_bytes;    /* device bytes of list headers + pool */
    int32_t chunk_slots;
    int32_t retries;          /* directory-overflow retries taken */
    int32_t n_colors;         /* colored path only */
    int32_t big_rows;         /* rows consolidated by the LDS (wave-per-row) kernel */
    int32_t graph_launches;   /* hipGraphLaunch calls the call made (batched assembly: 1 per call) */
    int32_t _pad;
} alls_stats;

/* ------------------------------------------------------------ context --- */
const char* alls_status_str(alls_status s);
/* Creates a context on `device`.  `hip_stream` is a hipStream_t, or NULL for
 * a stream the context creates and owns. */
alls_status alls_ctx_create(alls_ctx** ctx, int device, void* hip_stream);
void        alls_ctx_destroy(alls_ctx* ctx);
const char* alls_ctx_last_error(const alls_ctx* ctx);
alls_status alls_last_stats(const alls_ctx* ctx, alls_stats* out);
/* Number of visible GPUs (0 when none). */
int         alls_device_count(void);

/* ----------------------------------------------------------- assembly --- */
/* Plain A-LLS assembly: element integration, atomic insertion into row-wise
 * linked lists, and device consolidation into CSR.
 * Replaces, as one call:
 *   buildCSRStructure                   Acc:605-645 / T3:458-482 / Mix:488-512 / Hex:364-392 (host std::set pattern)
 *   assembleGlobalStiffnessKernelALLS   Acc:394-468 / T3:227-290 / Mix:248-312 (+ FixedArray, Hex:208-257)
 *   convertALLStoCSRKernel              Acc:471-501 / T3:292-314 / Mix:314-339 / Hex:259-278
 * and returns the values the reference's serial assembleGlobalStiffnessHost
 * (Acc:329-380) computes on that pattern.
 * The output is allocated (or `out`'s buffers reused when large enough) with
 * room for the list-entries bound, nnz <= entries x dpn^2 (cap_nnz), so that
 * the pipeline runs with one host synchronisation; when that bound passes
 * 2^31 (and opts->row_ptr_i64 is 0), or its allocation fails, the output is
 * sized exactly after a read-back of nnz.  opts->timing 0 records no stage
 * events (alls_last_stats then reports ms_total only). */
alls_status alls_assemble(alls_ctx* ctx, const alls_mesh* mesh, const alls_opts* opts, alls_csr* out);

/* alls_assemble without any device allocation.  No reference counterpart:
 * it is for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The list pool and
 * consolidation arrays live in a fixed device array of the library, 128 MiB
 * per GPU, held until the process ends; a mesh whose lists need more
 * returns ALLS_E_UNSUPPORTED (structured meshes of up to about 130k Q4
 * elements, 174k triangles, or 1,900 Hex20 elements in FP64 and 4,200 in
 * FP32-MFMA fit).  The mesh arrays must be device memory (mem =
 * ALLS_MEM_DEVICE, 16-byte aligned); opts->autotune must be 0.  `out` holds
 * the caller's device buffers: row_ptr (cap_rows entries, int64 when
 * row_ptr_is_i64 is set, else int32), col_ind and val (cap_nnz entries);
 * they are filled and n_rows/nnz set, never reallocated or freed.  If they
 * are too small the call returns ALLS_E_INVALID with out->n_rows and
 * out->nnz set to the sizes needed.  Every call on one GPU shares that
 * GPU's scratch: calls that overlap on one GPU (two contexts, two threads)
 * race.  Values are bitwise those of alls_assemble. */
alls_status alls_assemble_noalloc(alls_ctx* ctx, const alls_mesh* mesh, const alls_opts* opts, alls_csr* out);

/* Copies a CSR to host memory.  row_ptr is always returned as int64 and val
 * as double (float values widened exactly).  Any of the three destinations
 * may be NULL. */
alls_status alls_csr_download(alls_ctx* ctx, const alls_csr* csr, int64_t* row_ptr,
                              int32_t* col_ind, double* val);
void        alls_csr_free(alls_ctx* ctx, alls_csr* csr);

/* ------------------------------------------------- cached sparsity pattern */
/* Runs alls_assemble and keeps a device copy of the mesh, its CSR pattern and
 * an element -> CSR-slot scatter map, so later value-only assemblies skip
 * list construction.  The reference rebuilds everything for each run and
 * keeps its host pattern only (Acc:1014-1016). */
alls_status alls_pattern_create(alls_ctx* ctx, const alls_mesh* mesh, const alls_opts* opts,
                                alls_pattern** pat, alls_csr* out);
/* Value-only re-assembly on a cached pattern with a new per-element E field
 * (E_per_elem NULL => the E the pattern was built with).  No lists and no
 * atomics: in 2D one thread per node row recomputes its incident elements
 * and writes the row; for Hex20 the element blocks are staged in element
 * order and one wave per node row gathers its incident elements' blocks.
 * Summation runs in ascending element id, so repeated calls give values
 * bitwise equal to each other and to alls_assemble.  row_ptr/col_ind are
 * copied into `out` only when it does not already hold this pattern. */
alls_status alls_assemble_values(alls_ctx* ctx, const alls_pattern* pat, const double* E_per_elem,
                                 int32_t mem, alls_csr* out);
int64_t     alls_pattern_nnz(const alls_pattern* pat);
void        alls_pattern_free(alls_ctx* ctx, alls_pattern* pat);

/* ------------------------------------------------------- GPU coloring --- */
/* Element colouring through shared nodes on the GPU (Jones-Plassmann-Luby
 * with first-fit colours).  Replaces the host greedy colourers colorElements
 * (Acc:648-747, T3:484-523) and colorElementsHex20 (Hex:394-417). */
alls_status alls_color(alls_ctx* ctx, const alls_pattern* pat, alls_coloring** col);
/* Wraps a caller-provided colouring (host int32 [n_elems]); it is checked for
 * conflicts on the device. */
alls_status alls_coloring_from_host(alls_ctx* ctx, const alls_pattern* pat, const int32_t* colors,
                                    alls_coloring** col);
int32_t     alls_coloring_count(const alls_coloring* col);
alls_status alls_coloring_download(alls_ctx* ctx, const alls_coloring* col, int32_t* colors);
void        alls_coloring_free(alls_ctx* ctx, alls_coloring* col);
/* Coloured assembly on a cached pattern: one full-grid launch per colour with
 * plain (non-atomic) "+=" into binary-searched CSR slots.  Replaces
 * assembleGlobalStiffnessKernelColored[NoAtomic] (Acc:504-572, T3:316-363,
 * Mix:341-390, Hex:281-321), which run all colours at once in one block each. */
alls_status alls_assemble_colored(alls_ctx* ctx, const alls_pattern* pat, const alls_coloring* col,
                                  const double* E_per_elem, int32_t mem, alls_csr* out);

/* ------------------------------------------------------------ batching --- */
/* Assembles n meshes of one dimension in one pipeline.  The meshes are
 * packed into one node/element space (mesh k's node ids shifted by the nodes
 * of meshes 0..k-1), so each stage is ONE launch over all of them whatever
 * their sizes; the insertion kernel runs persistent (opts->grid_per_cu,
 * default 8 blocks per CU for 2D and 32 for Hex20); and the whole pipeline
 * runs as ONE hipGraph (into an output sized by the list entries bound, so
 * no host read-back splits it), captured on the first call and replayed
 * while the batch keeps its shape (sizes and material layout; new
 * coordinates or E values replay).
 * The result is the block-diagonal CSR of the batch: mesh k owns rows
 * [row_off[k], row_off[k+1]) and nonzeros [nnz_off[k], nnz_off[k+1]), and
 * its column indices are shifted by row_off[k].  row_off and nnz_off are
 * caller arrays of n+1 entries (host).  The meshes must share dim, mem, nu
 * and (2D) thickness; Q4, T3 and mixed 2D meshes may be batched together.
 * No reference counterpart: each reference program assembles one mesh per
 * call (Acc:750-838, T3:525, Mix:578, Hex:420). */
alls_status alls_assemble_batch(alls_ctx* ctx, int32_t n, const alls_mesh* meshes, const alls_opts* opts,
                                alls_csr* out, int64_t* row_off, int64_t* nnz_off);

/* Times the pipeline on `mesh` for each candidate launch setting (threads
 * per block and persistent grid of the insertion kernel, list-pool chunk
 * size), keeps the fastest for the mesh's element kind, precision and size
 * class (a power of two of the element count) in the context, and returns
 * it in *best (the other fields copied from *base).  Calls with
 * opts->autotune = 1 use the kept setting, tuning on their own mesh the
 * first time a kind and size class is seen.  Outputs do not depend on the
 * setting: every column is summed in element order. */
alls_status alls_autotune(alls_ctx* ctx, const alls_mesh* mesh, const alls_opts* base, alls_opts* best);

/* ------------------------------------------------------------- serving --- */
/* Continuous-batching assembly service on one GPU.  Requests are admitted
 * in arrival order as soon as a stream is idle and their HBM footprint fits
 * the budget, run concurrently on up to n_streams streams, and reuse cached
 * patterns when mesh_key repeats. */
/* When a request with a mesh key builds and caches its pattern.  COST (the
 * default): when the key's expected further repeats, times the share of
 * predicted hits that cached patterns have delivered before eviction, times
 * the saving of a cached repeat exceed the extra cost of the build.  The
 * expected repeats of a key seen k times are read from the server's own
 * history (keys seen k times that came back, per dimension); the costs are
 * per-element running averages the server measures per path and dimension.
 * EAGER: on the first request of every key. */
typedef enum { ALLS_CACHE_COST = 0, ALLS_CACHE_EAGER = 1 } alls_cache_policy;

typedef struct {
    int32_t device;
    int32_t n_streams;        /* concurrent in-flight requests; 0 => 3.  With 2 or more the last
                                 is an express lane for small requests; the others take the queue
                                 in order, one heavy request (over about 8 ms) at a time */
    int64_t hbm_budget_bytes; /* admission budget; 0 => 55% of free HBM, measured after the
                                 process's idle cached blocks are returned to the device */
    int32_t pattern_cache_entries;
    int32_t cache_policy;     /* alls_cache_policy */
} alls_server_config;

typedef struct {
    double queue_ms;          /* submit -> admitted */
    double run_ms;            /* admitted -> done (device work of this request) */
    double latency_ms;        /* submit -> done */
    int64_t nnz;
    int32_t cache_hit;        /* 1 => value-only assembly on a cached pattern, 2 => this request
                                 built and cached the pattern, 0 => plain assembly */
    int32_t status;           /* alls_status of the request */
    int32_t batch_size;       /* requests assembled in the same packed launch (1: alone) */
    int32_t reserved;
} alls_request_stats;

/* What a server has done so far (alls_server_counts).  Index 0 counts 2D
 * requests, index 1 Hex20. */
typedef struct {
    int64_t completed;        /* requests finished (any path, any status) */
    int64_t batched;          /* of them, assembled in a packed batch with others */
    int64_t passed;           /* admitted ahead of an older queued request (express lane) */
    int64_t builds[2];        /* patterns built and cached */
    int64_t hits[2];          /* value-only assemblies on a cached pattern */
    int64_t evictions[2];     /* cached patterns dropped to make room (or replaced under their key) */
    int64_t evicted_unused[2];/* of them, dropped before any hit */
} alls_server_counts;

alls_status alls_server_create(const alls_server_config* cfg, alls_server** srv);
/* Copies the server's counters into *out. */
alls_status alls_server_counters(alls_server* srv, alls_server_counts* out);
/* Enqueues a request; the mesh arrays must stay valid until it completes.
 * mesh_key != 0 enables the pattern cache for this mesh.  A key promises the
 * same node coordinates and connectivity on every request that carries it:
 * only E (scalar or per element) may change.  A cached pattern serves a
 * request only if precision, node and element counts, dim, nodes_per_elem,
 * uniform_type, the presence of elem_type, nu and (2D) thickness all match;
 * otherwise the request runs the plain path (or rebuilds the pattern).
 * Returns ALLS_E_BUSY if the request alone exceeds the whole budget. */
alls_status alls_server_submit(alls_server* srv, const alls_mesh* mesh, uint64_t mesh_key,
                               const alls_opts* opts, int64_t* ticket);
/* Blocks until the request completes.  Its CSR moves into *out (the caller
 * frees it with alls_server_release, before alls_server_destroy: outputs live
 * in the server's HBM reservation). */
alls_status alls_server_wait(alls_server* srv, int64_t ticket, alls_csr* out, alls_request_stats* st);
void        alls_server_release(alls_server* srv, alls_csr* csr);
void        alls_server_destroy(alls_server* srv);

/* --------------------------------------------------- synthetic meshes --- */
/* Host-side generators of the benchmark meshes (own counter-based RNG).
 * They replace generateMesh (Acc:575-603), MeshGenerator (T3:366-456,
 * Mix:393-486) and generateMeshHex20 (Hex:324-362), adding node jitter,
 * random diagonals, node-id permutation and lognormal E. */
typedef enum {
    ALLS_GEN_Q4_GRID = 0,     /* nx*ny Q4 cells on the unit square */
    ALLS_GEN_T3_GRID = 1,     /* every cell split into 2 T3 */
    ALLS_GEN_MIXED_GRID = 2,  /* split_fraction of cells split into 2 T3, rest Q4 */
    ALLS_GEN_HEX20_BRICK = 3  /* nx*ny*nz Hex20 on the unit cube, compact node numbering */
} alls_gen_kind;

typedef struct {
    int32_t kind;             /* alls_gen_kind */
    int32_t nx, ny, nz;
    double jitter;            /* interior-node jitter, fraction of the cell size (uniform +-) */
    uint64_t seed;
    int32_t random_diagonal;  /* T3/mixed: 1 => random split diagonal, 0 => n0-n2 as the reference */
    int32_t permute_nodes;    /* 1 => random permutation of node ids */
    double split_fraction;    /* mixed: probability that a cell is split */
    double E;                 /* scalar E (used when E_sigma == 0) */
    double E_sigma;           /* > 0 => per-element E = exp(E_sigma * N(0,1)) */
    double nu;
    double thickness;
} alls_gen_spec;

typedef struct {
    alls_mesh mesh;           /* host arrays owned by the library */
    int32_t* structured_color;/* [n_elems] cell-parity colouring (4 colours in 2D, 8 in 3D) */
    int32_t n_structured_colors;
    int32_t _pad;
} alls_host_mesh;

alls_status alls_gen_mesh(const alls_gen_spec* spec, alls_host_mesh* out);
void        alls_host_mesh_free(alls_host_mesh* m);

/* ------------------------------------------------------------- files --- */
/* Writes a CSR as a Matrix Market coordinate file (real general, 1-based,
 * every stored entry, values printed with 17 significant digits so they
 * read back bit-exact).  The reference writes only CSV timing tables
 * (Perf:1139, Mix:838). */
alls_status alls_csr_write_mtx(alls_ctx* ctx, const alls_csr* csr, const char* path);
/* Binary mesh file: a 96-byte header (magic "ALLSMSH1", dim, nodes_per_elem,
 * uniform_type, flags, n_nodes, n_elems, E, nu, thickness), then coords,
 * conn, elem_type (flag 1) and E_per_elem (flag 2) as stored in memory.
 * Save takes a host mesh; load returns one owned by the library (free with
 * alls_host_mesh_free). */
alls_status alls_mesh_save(const alls_mesh* mesh, const char* path);
alls_status alls_mesh_load(const char* path, alls_host_mesh* out);

/* ----------------------------------------------------- device rates --- */
/* Rates measured on the context's device, the measured side of the
 * roofline figures (the vendor side: 8 TB/s, 78.6 / 157.3 TFLOP/s).  No
 * reference counterpart: the reference quotes no peaks.  Takes about a
 * second and 2 GiB of device memory, freed on return. */
typedef struct {
    double copy_GBs;          /* 1 GiB 16-byte streaming copy, bytes read + written per second */
    double mfma_f64_TFLOPs;   /* v_mfma_f64_16x16x4f64 on register operands, every CU */
    double mfma_f32_TFLOPs;   /* v_mfma_f32_16x16x4f32, the same way */
    double atomic_f64_Gops;   /* FP64 atomic adds (no return) over 16M distinct doubles */
} alls_peaks;

alls_status alls_device_peaks(alls_ctx* ctx, alls_peaks* out);

#ifdef __cplusplus
}
#endif
#endif /* ALLS_H */
