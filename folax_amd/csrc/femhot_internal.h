// Library-internal context and helpers (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <list>
#include <string>
#include <vector>

#include "../../include/femhot.h"
#include "femhot_element.h"

namespace femhot {

void set_error(const std::string &msg);
const std::string &get_error();

#define FH_CHECK_ARG(cond, code, msg)                                    \
  do {                                                                   \
    if (!(cond)) {                                                       \
      ::femhot::set_error(std::string(__func__) + ": " + (msg));         \
      return (code);                                                     \
    }                                                                    \
  } while (0)

#define FH_HIP(call)                                                                        \
  do {                                                                                      \
    hipError_t _e = (call);                                                                 \
    if (_e != hipSuccess) {                                                                 \
      ::femhot::set_error(std::string(__func__) + ": " #call " -> " + hipGetErrorString(_e)); \
      return FEMHOT_ERR_HIP;                                                                \
    }                                                                                       \
  } while (0)

inline int nodes_per_elem(int et) { return et == HEX8 ? 8 : 4; }
inline int elem_dim(int et) { return et == QUAD4 ? 2 : 3; }

// Device-resident chunk plan for the LDS-staged scatter (see femhot_chunk.hip).
struct ChunkPlan {
  int epc = 0;             // max elements per chunk the plan was built for
  int n_chunks = 0;
  int max_nodes = 0;       // max local nodes over chunks
  int max_elems = 0;
  long long n_shared = 0;  // nodes touched by more than one chunk
  int *chunk_elem_off = nullptr;  // (n_chunks+1) offsets into perm / lconn
  int *chunk_node_off = nullptr;  // (n_chunks+1) offsets into chunk_nodes
  int *chunk_seg = nullptr;       // (n_chunks) segment id
  int *perm = nullptr;            // (Ne) original element id at each position
  bool perm_identity = false;     // chunks are runs of the caller's element order (perm[i] == i)
  uint16_t *lconn = nullptr;      // (Ne * n) chunk-local node indices, position order
  int *chunk_nodes = nullptr;     // global node id of each chunk node
  float *chunk_xyz32 = nullptr;   // per chunk node: (x, y, z, 0) relative to the chunk's first node
                                  // (16-byte records for the LDS-DMA kernel)
  float *chunk_xyz12 = nullptr;   // per chunk node: (x, y, z), 12-byte records for the chunk kernel
  double *chunk_xyz64 = nullptr;  // per chunk node: (x, y, z), 24-byte records
  long long n_zero = 0;           // nodes in no element (their outputs are cleared separately)
  int *zero_nodes = nullptr;
  // deterministic reduction of shared nodes: each (chunk, shared node) pair owns a partial slot
  long long n_slots = 0;
  int *node_slot = nullptr;       // per chunk node: partial slot, -1 when the node is exclusive
  int *sh_nodes = nullptr;        // (n_shared) shared node ids, ascending
  int *red_off = nullptr;         // (n_shared+1) prefix sums of the shared nodes' chunk counts
  int *seg_chunk_off = nullptr;   // (nseg+1) chunks of each segment (chunks are segment-sorted)
  // node-centric gather inside a chunk: for each chunk node, the (element position * n + corner)
  // entries touching it, sorted by element position; per chunk the list has n * elements entries
  uint16_t *inc = nullptr;        // (Ne * n), chunk by chunk
  uint16_t *inc_start = nullptr;  // (chunk nodes) start of each node's entries within its chunk
  uint16_t *rank = nullptr;       // (Ne * n) position of corner (element position, a) in its chunk's list
  unsigned long long last_use = 0;  // plan cache: least recently used plan is evicted first
};

// Chunk plans kept per context (one per elements-per-chunk in use).
constexpr size_t kMaxChunkPlans = 4;

// Device scratch of the chunk path (partials and per-chunk energies), grown on demand.
struct ChunkWork {
  size_t bytes = 0;
  void *ptr = nullptr;
};

struct Ctx {
  femhot_mesh_desc desc{};
  int et = 0, ph = 0, ngp = 0, n = 0, dim = 0, ndpn = 0;
  long long Nn = 0, Ne = 0, Ndof = 0;
  bool repeated_nodes = false;  // some element lists a node twice (collapsed element)
  PhysParams pp{};
  int nseg = 1;
  // host copies
  std::vector<double> h_coords;
  std::vector<int> h_conn;
  std::vector<long long> h_dir_idx;
  std::vector<double> h_dir_val;
  std::vector<long long> h_seg_off;
  // device mesh
  double *d_coords64 = nullptr;  // (Nn*3)
  float *d_coords32 = nullptr;   // (Nn*3)
  int *d_conn = nullptr;         // (Ne*n)
  uint8_t *d_dmask = nullptr;    // (Ndof) 1 = Dirichlet
  long long *d_seg_off = nullptr;  // (nseg+1)
  // lazily built caches
  bool have_pattern = false;
  std::vector<long long> h_row_ptr;  // (Ndof+1)
  std::vector<int> h_col_idx;
  long long nnz = 0;
  long long *d_row_ptr = nullptr;
  int *d_col_idx = nullptr;
  uint8_t *d_slot = nullptr;  // (Ne*n*n) position of node b in node a's sorted neighbour list
  int *d_diag_off = nullptr;  // (Ndof) offset of the diagonal entry within each CSR row
  int *d_ne_off = nullptr;    // (Nn+1) node -> incident (element, local index) entries
  int *d_ne_list = nullptr;   // (Ne*n) element * n + local index, element order within a node
  bool have_colour = false;
  std::vector<int> h_colour;
  int n_colours = 0;
  std::vector<int> h_colour_off;  // (n_colours+1)
  int *d_colour_elems = nullptr;  // elements sorted by colour (stable)
  std::list<ChunkPlan> plans;     // built on demand, cached per elements-per-chunk (stable addresses)
  unsigned long long plan_tick = 0;
  ChunkWork chunk_ws;
  ChunkWork newton_ws;            // femhot_newton: tangent values, residual, rhs, step, Dirichlet values
  bool newton_dval_ready = false;
  // kernel timing hook: events recorded around the dominant kernel of each call while enabled
  bool timing = false;
  std::vector<hipEvent_t> tev_pool;
  size_t tev_used = 0;
  // solver workspace
  size_t ws_bytes = 0;
  void *d_ws = nullptr;
  double *h_pinned = nullptr;  // pinned host scratch (1024 doubles: the solver's reduction partials)
};

int ensure_pattern(Ctx *c);
int ensure_colouring(Ctx *c);
// The chunk plan of `epc` elements per chunk, built (and cached) on first use; refused while `s`
// is being captured into a graph.
int ensure_chunks(Ctx *c, int epc, hipStream_t s, const ChunkPlan **out);
// The cached plan of `epc` elements per chunk, or null when it is not built (never builds).
const ChunkPlan *find_chunks(Ctx *c, int epc);
void free_chunk_plans(Ctx *c);
// Resolves, once per device with that device current, the address of the chunk path's fixed device
// scratch (the allocation-free entry point's); called by femhot_ctx_create.
int init_fixed_scratch();
int ensure_workspace(Ctx *c, size_t bytes);

// Records an event on `s` when timing is enabled (pairs: before/after the dominant kernel).
int timing_mark(Ctx *c, hipStream_t s);

}  // namespace femhot

struct femhot_ctx {
  femhot::Ctx c;
};
