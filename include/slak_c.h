/* C-ABI of the MI355X Catmull-Clark library (libslak.so).
 *
 * Plain pointers and sizes only. Mesh pointers are device pointers. Every
 * entry point returns a status code (SLAK_OK = 0); slak_last_error() gives the
 * message of the calling thread's last failure. The C++ headers
 * LAKCatmullClark.h / CatmullClark.h turn a failure into a throw of that
 * message as `const char*`, the reference's error convention
 * (LibInclude/cuda_host_helpers.h:34-41).
 */
#ifndef SLAK_C_H
#define SLAK_C_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  SLAK_OK = 0,
  SLAK_ERR_INVALID_ARGUMENT = 1,
  SLAK_ERR_INVALID_MESH = 2,
  SLAK_ERR_HIP = 3,
  SLAK_ERR_OUT_OF_MEMORY = 4,
  SLAK_ERR_INTERNAL = 5,
  SLAK_ERR_REJECTED = 6
};

/* MeshInfo::MeshType (LibInclude/LAKCatmullClark.h:98) */
enum { SLAK_TRI = 0, SLAK_QUAD = 1, SLAK_POLY = 2 };

/* LAKCatmullClark (sparse linear algebra) or CatmullClark (specialised). */
enum { SLAK_SCHEME_LAK = 0, SLAK_SCHEME_SLAK = 1 };

/* MeshInfo<int,int,int,float> (LibInclude/LAKCatmullClark.h:95-156,
 * LibInclude/CatmullClark.h:69-130): face-vertex incidence matrix with one CSR
 * row per face; vals[k] is the 1-based position of the corner in its face;
 * verts holds float4 (x, y, z, w = 1) per vertex. All pointers on the device. */
typedef struct slak_mesh {
  int32_t type;
  int32_t is_reduced;
  uint64_t nfaces;
  uint64_t nverts;
  uint64_t nnz;
  uint64_t max_face_order;
  int32_t* ptr;
  int32_t* ids;
  int32_t* vals;
  float* verts;
} slak_mesh;

/* Per-call report: replaces the reference's stdout banner
 * (LAKCatmullClarkImpl.cu:560-570, CatmullClarkImpl.cu:504-514). */
typedef struct slak_stats {
  double wall_ms;       /* host wall clock of the whole call, inputs resident */
  uint64_t peak_bytes;  /* peak device bytes held by the library's pool */
  int32_t levels;       /* levels refined */
  int32_t host_syncs;   /* blocking device->host round trips */
  uint64_t out_faces;
  uint64_t out_verts;
} slak_stats;

/* LAKCatmullClark::subdivideIteratively / CatmullClark::subdivideIteratively
 * (LibInclude/LAKCatmullClark.h:190-191, LibInclude/CatmullClark.h:163-164;
 * LibSource/device/LAKCatmullClarkImpl.cu:512-571,
 * LibSource/device/CatmullClarkImpl.cu:457-515).
 * A non-quad input always gets one level; quad levels follow up to
 * target_level. On success rmesh receives newly allocated device buffers owned
 * by the caller (free them with slak_free): LAK form has ptr/ids/vals
 * (ptr[i] = 4i, vals cyclic 1..4, is_reduced = 0), SLAK form ids only
 * (is_reduced = 1). A quad input with target_level 0 leaves rmesh untouched.
 * The input is never modified. stats and stream may be NULL. Synchronous. */
int slak_subdivide(int32_t scheme, const slak_mesh* cmesh, slak_mesh* rmesh, int32_t target_level,
                   slak_stats* stats, void* stream);

/* Frees a buffer returned by the library (replaces cudaFree in
 * MeshInfo::freeAndReset, LibInclude/LAKCatmullClark.h:120-127). It waits for
 * the legacy default stream (and so for the blocking streams), not for the
 * whole device: work of the caller's own non-blocking streams that still
 * reads the buffer must be synchronized first. Also accepts slak_malloc
 * pointers. */
int slak_free(void* device_ptr);

/* Message of this thread's last failure ("" if none). */
const char* slak_last_error(void);

/* Library version, e.g. 100 for 0.1.0. */
int32_t slak_version(void);

/* ---- device memory helpers for hosts without a GPU runtime binding ---- */
int slak_malloc(void** device_ptr, uint64_t bytes);
int slak_memcpy_h2d(void* dst_device, const void* src_host, uint64_t bytes);
int slak_memcpy_d2h(void* dst_host, const void* src_device, uint64_t bytes);
int slak_memcpy_d2d(void* dst_device, const void* src_device, uint64_t bytes);
int slak_device_synchronize(void);
/* Returns cached pool blocks to the device. */
int slak_pool_trim(void);
/* Bytes the pool holds in use / cached. */
int slak_pool_stats(uint64_t* in_use, uint64_t* cached, uint64_t* peak);

/* ---- per-stage timing (TimeProfileIndividual, LibInclude/profiling.h:182-200) ----
 * slak_profile_enable(1) clears the record and brackets every later kernel
 * launch with a HIP event pair on its stream; slak_profile_report writes
 * {"stages": [{"name", "launches", "total_ms", "bytes"}]} as JSON text, where
 * bytes are the stage's algorithmic (compulsory) bytes summed over launches. */
int slak_profile_enable(int32_t on);
int slak_profile_report(char* buf, uint64_t len);
/* HBM reference: `iters` float4 stream copies of `bytes` (k_copy_probe);
 * *gbps = (read + write bytes) / time. Also the calibration kernel for the
 * FETCH_SIZE / WRITE_SIZE counters. */
int slak_copy_probe(uint64_t bytes, int32_t iters, double* gbps);

/* Order-sensitive 64-bit digest of `bytes` bytes of device memory: the sum
 * mod 2^64 over the 32-bit words w_i (the last zero-padded) of
 * splitmix64(w_i + (i + 1) * 0x9E3779B97F4A7C15). Compares device results
 * (a batched mesh with the same mesh refined alone) without copying them to
 * the host. No counterpart in the reference. Synchronous. */
int slak_digest(const void* device_ptr, uint64_t bytes, uint64_t* digest, void* stream);

/* ---- sparse operators of the LAK method (LibInclude/math/bhblas.h:30-51) ---- */

/* Mapped SpGEMM of one level (spgemm_mapped, LibInclude/device/math/spgemm.cuh:102-309,
 * followed by compress, LibInclude/device/LAKCatmullClarkKernels.cuh:217-252):
 * F column v lists w with value f+1 for the half-edge v->w of face f; D column
 * v lists w with value f+1 for w->v; E column c lists r < c with value -id
 * (interior) or +id (boundary), ids 1-based in (c, r) order. Rows ascending.
 * Output arrays are newly allocated device buffers (free with slak_free):
 * ptr arrays have nverts+1 entries. *nextern receives the boundary edge count.
 * The input mesh is a level-input mesh as for slak_subdivide. */
int slak_spgemm_mapped(const slak_mesh* cmesh, int32_t** F_ptr, int32_t** F_ids, int32_t** F_vals,
                       int32_t** D_ptr, int32_t** D_ids, int32_t** D_vals, int32_t** E_ptr, int32_t** E_ids,
                       int32_t** E_vals, uint64_t* nedges, uint64_t* nextern, void* stream);

/* ---- reduced / unreduced forms of homogeneous (TRI, QUAD) meshes ---- */

/* GPU::CSCMesh::reduce (include/Mesh.h:305-334): each face's (ids, vals) is
 * reordered in place by vals, so that ids are in corner-position order, and
 * is_reduced is set. A mesh already reduced is left as it is. POLY meshes are
 * rejected ("mesh type not homogeneous"), as is an unreduced mesh without
 * vals. Synchronous; device pointers in *mesh. */
int slak_mesh_reduce(slak_mesh* mesh, void* stream);
/* GPU::CSCMesh::unreduce (include/Mesh.h:336-377): a missing ptr is allocated
 * with stride offsets and missing vals with 1-based corner positions (both
 * freed with slak_free); each face's (ids, vals) is then reordered in place by
 * ids and is_reduced is cleared. The reference fills missing vals with
 * (c - 1) mod order (-1, 0, 1, ...), which no LAK kernel accepts as
 * positions; the fill here is the position the rest of this API uses. */
int slak_mesh_unreduce(slak_mesh* mesh, void* stream);

/* Exclusive prefix sum with total: out[i] = sum(in[0..i)), out[n] = sum(in)
 * (CubFunctions::scan_exclusive, LibInclude/device/cub_wrappers.cuh:30-43,
 * which also writes n+1 values). Device pointers; out holds n+1 int32. */
int slak_scan_exclusive(const int32_t* in, int32_t* out, uint64_t n, void* stream);

/* slak_scan_exclusive without any allocation. No counterpart in the
 * reference: for callers that cannot allocate device memory (code between
 * allocations they control, serving loops, harnesses that forbid hipMalloc).
 * Its flags word and tile partials live in a 1 MiB static device array per
 * GPU, so n is limited to 268,367,872 items (SLAK_ERR_INVALID_ARGUMENT
 * beyond). Every call on a GPU shares that array: calls that overlap on one
 * GPU (two host threads, or two streams) race and must be serialized by the
 * caller. Synchronous, like slak_scan_exclusive. */
int slak_scan_exclusive_noalloc(const int32_t* in, int32_t* out, uint64_t n, void* stream);

/* ---- batched subdivision (no counterpart in the reference, which refines
 * one mesh per call) ----
 * slak_batch_create validates the meshes (device pointers, kept by pointer:
 * they must stay valid and unchanged in topology until destroy), runs the
 * first level's analysis for all of them (two blocking round trips) and
 * allocates every buffer of the whole chain, sized in closed form.
 * slak_batch_run recomputes the full chain from the current inputs (positions
 * may change between runs) with one launch per kernel per level over the union
 * of the meshes still active at that level; use_graph = 1 replays it as one
 * hipGraph captured on the first run. Synchronous.
 * slak_batch_result gives a view of mesh `index`'s refined mesh in the form
 * slak_subdivide returns; the buffers belong to the batch (do not free). */
typedef struct slak_batch slak_batch;

typedef struct slak_batch_info {
  int32_t nmeshes;
  int32_t max_levels;
  int32_t host_syncs;
  int32_t reserved;
  uint64_t device_bytes;
  uint64_t out_faces;
  uint64_t out_verts;
} slak_batch_info;

int slak_batch_create(int32_t scheme, int32_t nmeshes, const slak_mesh* meshes, const int32_t* target_levels,
                      slak_batch** batch, void* stream);
int slak_batch_run(slak_batch* batch, int32_t use_graph);
int slak_batch_result(slak_batch* batch, int32_t index, slak_mesh* rmesh);
int slak_batch_info_get(slak_batch* batch, slak_batch_info* info);
int slak_batch_destroy(slak_batch* batch);

/* ---- serving (no counterpart in the reference) ----
 * A request loop with continuous batching: submitted requests are admitted
 * (FIFO) while the sum of their projected device footprints stays within
 * memory_budget_bytes, at most max_batch per step. A step refines each of its
 * requests to the target level in the batched kernels (one launch per kernel
 * per level over the union of the requests still active at that level; the
 * first level's analysis once, later levels' structure in closed form, the
 * last level from the quad mesh two levels back). Up to `slots` steps run at
 * once, each on its own stream, so requests that arrive while others are on
 * the GPU start beside them as soon as a slot and memory are free.
 * Intermediate levels stay resident on the device. The submitted mesh must
 * stay valid until the request completes. Results are identical to
 * slak_subdivide of the same mesh. */
typedef struct slak_server slak_server;

typedef struct slak_server_config {
  int32_t scheme;
  int32_t device;
  int32_t max_batch;              /* requests per step; <= 0: 1024 */
  int32_t slots;                  /* concurrent steps, 0..8; 0: 2 */
  uint64_t memory_budget_bytes;   /* 0: 64 GiB */
} slak_server_config;

typedef struct slak_request_stats {
  int32_t status;
  int32_t levels;
  int32_t steps;       /* batched steps the request took part in (1) */
  int32_t reserved;
  double submit_us;    /* times relative to the server start */
  double admit_us;
  double done_us;      /* result resident on the device */
  double latency_us;   /* done - submit */
  uint64_t out_faces;
} slak_request_stats;

typedef struct slak_server_stats {
  uint64_t steps;
  uint64_t completed;
  uint64_t pending;
  uint64_t active;
  uint64_t inflight_bytes;
  uint64_t faces_done;
  double busy_us;      /* host time inside batched steps */
  int32_t max_batch;   /* largest batch seen */
  int32_t reserved;
} slak_server_stats;

int slak_server_start(const slak_server_config* cfg, slak_server** server);
/* Returns immediately with a request id; SLAK_ERR_REJECTED when the request
 * alone exceeds the memory budget or the server is stopping. */
int slak_server_submit(slak_server* server, const slak_mesh* mesh, int32_t target_level, uint64_t* request_id);
/* Waits (timeout_ms < 0: forever) for a request; on success *result receives
 * its refined mesh, owned by the caller (slak_free), in the slak_subdivide form. */
int slak_server_wait(slak_server* server, uint64_t request_id, int32_t timeout_ms, slak_mesh* result,
                     slak_request_stats* stats);
int slak_server_get_stats(slak_server* server, slak_server_stats* stats);
/* Finishes every submitted request, then stops; uncollected results are freed. */
int slak_server_stop(slak_server* server);

#ifdef __cplusplus
}
#endif

#endif /* SLAK_C_H */
