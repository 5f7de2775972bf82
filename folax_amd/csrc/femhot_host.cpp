// Host side of the femhot library: context creation, mesh validation, CSR pattern, slot table,
// colouring, chunk plans and error reporting.  Built with hipcc as host code into libfemhot.so.
//
// Reference behaviour followed:
//   fe_loss.py:34-55    Dirichlet dof numbering ndpn*node + comp
//   fe_loss.py:57-122   Initialize: element type, Gauss rule defaults (num_gp), dofs per node
//   fe_loss.py:178-184, 313-316 + fe_solver.py:71-72
//                       COO of element blocks -> scipy csr_array (summed duplicates, sorted
//                       columns, explicit zeros kept): the pattern built here
//   mesh.py:123-142     orientation fix (swap local nodes 0 and 1 where detJ < 0)
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "femhot_chunk.h"
#include "femhot_internal.h"

namespace femhot {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }
const std::string &get_error() { return g_err; }

static int default_num_gp(int et) { return et == TET4 ? 1 : 2; }

// Build node adjacency (sorted unique neighbour lists incl. self) and the CSR pattern + slot table.
int ensure_pattern(Ctx *c) {
  if (c->have_pattern) return FEMHOT_OK;
  const int n = c->n, P = c->ndpn;
  const long long Nn = c->Nn, Ne = c->Ne;
  if (Ne * n > 2147483647LL) {
    set_error("ensure_pattern: more than 2^31 - 1 element corners");
    return FEMHOT_ERR_SHAPE;
  }
  std::vector<long long> cnt(Nn + 1, 0);
  for (long long e = 0; e < Ne; ++e)
    for (int a = 0; a < n; ++a) cnt[c->h_conn[e * n + a] + 1]++;
  for (long long i = 0; i < Nn; ++i) cnt[i + 1] += cnt[i];
  std::vector<long long> n2e(cnt[Nn]);
  // node -> (element, local index) incidence of the row-owner tangent assembly, element order
  std::vector<int> ne_list(cnt[Nn]), ne_off(Nn + 1);
  {
    std::vector<long long> pos(cnt.begin(), cnt.end() - 1);
    for (long long e = 0; e < Ne; ++e)
      for (int a = 0; a < n; ++a) {
        const long long k = pos[c->h_conn[e * n + a]]++;
        n2e[k] = e;
        ne_list[k] = (int)(e * n + a);
      }
    for (long long i = 0; i <= Nn; ++i) ne_off[i] = (int)cnt[i];
  }
  std::vector<long long> adj_off(Nn + 1, 0);
  std::vector<int> adj;
  adj.reserve((size_t)Nn * 16);
  std::vector<int> tmp;
  for (long long i = 0; i < Nn; ++i) {
    tmp.clear();
    for (long long k = cnt[i]; k < cnt[i + 1]; ++k) {
      const long long e = n2e[k];
      for (int b = 0; b < n; ++b) tmp.push_back(c->h_conn[e * n + b]);
    }
    std::sort(tmp.begin(), tmp.end());
    tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
    adj.insert(adj.end(), tmp.begin(), tmp.end());
    adj_off[i + 1] = (long long)adj.size();
  }
  // CSR rows: node-major, P rows per node, each with P*deg columns (neighbour m, comp d) sorted
  c->h_row_ptr.assign(c->Ndof + 1, 0);
  long long base = 0;
  for (long long i = 0; i < Nn; ++i) {
    const long long deg = adj_off[i + 1] - adj_off[i];
    for (int cc = 0; cc < P; ++cc) c->h_row_ptr[i * P + cc] = base + (long long)cc * P * deg;
    base += (long long)P * P * deg;
  }
  c->h_row_ptr[c->Ndof] = base;
  c->nnz = base;
  if (base > 2147483647LL * 8) {
    set_error("ensure_pattern: tangent too large");
    return FEMHOT_ERR_SHAPE;
  }
  c->h_col_idx.resize(base);
  for (long long i = 0; i < Nn; ++i) {
    const long long deg = adj_off[i + 1] - adj_off[i];
    for (int cc = 0; cc < P; ++cc) {
      long long r = c->h_row_ptr[i * P + cc];
      for (long long k = 0; k < deg; ++k)
        for (int d = 0; d < P; ++d) c->h_col_idx[r++] = adj[adj_off[i] + k] * P + d;
    }
  }
  // diagonal of each row: offset within the row of column (i, c) for row (i, c), -1 for an empty row
  std::vector<int> diag_off(c->Ndof);
  for (long long i = 0; i < Nn; ++i) {
    const int *beg = adj.data() + adj_off[i], *end = adj.data() + adj_off[i + 1];
    const int *at = std::lower_bound(beg, end, (int)i);  // a node in some element is its own neighbour
    const bool has = at != end && *at == (int)i;         // (a node in no element has an empty row)
    for (int cc = 0; cc < P; ++cc) diag_off[i * P + cc] = has ? (int)((at - beg) * P + cc) : -1;
  }
  // slot table: position of node b in the sorted neighbour list of node a, per element
  std::vector<uint8_t> slot((size_t)Ne * n * n);
  for (long long e = 0; e < Ne; ++e)
    for (int a = 0; a < n; ++a) {
      const int na = c->h_conn[e * n + a];
      const int *beg = adj.data() + adj_off[na], *end = adj.data() + adj_off[na + 1];
      for (int b = 0; b < n; ++b) {
        const int nb = c->h_conn[e * n + b];
        const long long k = std::lower_bound(beg, end, nb) - beg;
        if (k > 255) {
          set_error("ensure_pattern: node with more than 256 neighbours is not supported");
          return FEMHOT_ERR_UNSUPPORTED;
        }
        slot[((size_t)e * n + a) * n + b] = (uint8_t)k;
      }
    }
  FH_HIP(hipMalloc(&c->d_row_ptr, sizeof(long long) * (c->Ndof + 1)));
  FH_HIP(hipMalloc(&c->d_col_idx, sizeof(int) * std::max<long long>(1, base)));
  FH_HIP(hipMalloc(&c->d_slot, slot.size()));
  FH_HIP(hipMemcpy(c->d_row_ptr, c->h_row_ptr.data(), sizeof(long long) * (c->Ndof + 1), hipMemcpyHostToDevice));
  if (base) FH_HIP(hipMemcpy(c->d_col_idx, c->h_col_idx.data(), sizeof(int) * base, hipMemcpyHostToDevice));
  FH_HIP(hipMemcpy(c->d_slot, slot.data(), slot.size(), hipMemcpyHostToDevice));
  FH_HIP(hipMalloc(&c->d_diag_off, sizeof(int) * std::max<long long>(1, c->Ndof)));
  if (c->Ndof) FH_HIP(hipMemcpy(c->d_diag_off, diag_off.data(), sizeof(int) * c->Ndof, hipMemcpyHostToDevice));
  FH_HIP(hipMalloc(&c->d_ne_off, sizeof(int) * (Nn + 1)));
  FH_HIP(hipMalloc(&c->d_ne_list, sizeof(int) * std::max<size_t>(1, ne_list.size())));
  FH_HIP(hipMemcpy(c->d_ne_off, ne_off.data(), sizeof(int) * (Nn + 1), hipMemcpyHostToDevice));
  if (!ne_list.empty())
    FH_HIP(hipMemcpy(c->d_ne_list, ne_list.data(), sizeof(int) * ne_list.size(), hipMemcpyHostToDevice));
  c->have_pattern = true;
  return FEMHOT_OK;
}

// Deterministic greedy first-fit colouring in element order over the shared-node graph.
int ensure_colouring(Ctx *c) {
  if (c->have_colour) return FEMHOT_OK;
  const int n = c->n;
  constexpr int W = 4;  // up to 256 colours
  std::vector<uint64_t> used((size_t)c->Nn * W, 0);
  c->h_colour.assign(c->Ne, 0);
  int ncol = 0;
  for (long long e = 0; e < c->Ne; ++e) {
    uint64_t m[W] = {0, 0, 0, 0};
    for (int a = 0; a < n; ++a) {
      const uint64_t *u = &used[(size_t)c->h_conn[e * n + a] * W];
      for (int w = 0; w < W; ++w) m[w] |= u[w];
    }
    int col = -1;
    for (int w = 0; w < W && col < 0; ++w)
      if (~m[w]) col = w * 64 + __builtin_ctzll(~m[w]);
    if (col < 0) {
      set_error("ensure_colouring: more than 256 colours needed");
      return FEMHOT_ERR_UNSUPPORTED;
    }
    c->h_colour[e] = col;
    ncol = std::max(ncol, col + 1);
    for (int a = 0; a < n; ++a) used[(size_t)c->h_conn[e * n + a] * W + col / 64] |= (1ULL << (col % 64));
  }
  c->n_colours = ncol;
  c->h_colour_off.assign(ncol + 1, 0);
  for (long long e = 0; e < c->Ne; ++e) c->h_colour_off[c->h_colour[e] + 1]++;
  for (int k = 0; k < ncol; ++k) c->h_colour_off[k + 1] += c->h_colour_off[k];
  std::vector<int> order(c->Ne);
  {
    std::vector<int> pos(c->h_colour_off.begin(), c->h_colour_off.end() - 1);
    for (long long e = 0; e < c->Ne; ++e) order[pos[c->h_colour[e]]++] = (int)e;
  }
  FH_HIP(hipMalloc(&c->d_colour_elems, sizeof(int) * std::max<long long>(1, c->Ne)));
  FH_HIP(hipMemcpy(c->d_colour_elems, order.data(), sizeof(int) * c->Ne, hipMemcpyHostToDevice));
  c->have_colour = true;
  return FEMHOT_OK;
}

static inline uint64_t spread3(uint64_t x) {  // 21 bits -> every third bit
  x &= 0x1fffff;
  x = (x | x << 32) & 0x1f00000000ffffULL;
  x = (x | x << 16) & 0x1f0000ff0000ffULL;
  x = (x | x << 8) & 0x100f00f00f00f00fULL;
  x = (x | x << 4) & 0x10c30c30c30c30c3ULL;
  x = (x | x << 2) & 0x1249249249249249ULL;
  return x;
}

static void free_chunks(ChunkPlan &p) {
  (void)hipFree(p.chunk_elem_off);
  (void)hipFree(p.chunk_node_off);
  (void)hipFree(p.chunk_seg);
  (void)hipFree(p.perm);
  (void)hipFree(p.lconn);
  (void)hipFree(p.chunk_nodes);
  (void)hipFree(p.chunk_xyz32);
  (void)hipFree(p.chunk_xyz12);
  (void)hipFree(p.chunk_xyz64);
  (void)hipFree(p.zero_nodes);
  (void)hipFree(p.node_slot);
  (void)hipFree(p.sh_nodes);
  (void)hipFree(p.red_off);
  (void)hipFree(p.seg_chunk_off);
  (void)hipFree(p.inc);
  (void)hipFree(p.inc_start);
  (void)hipFree(p.rank);
  p = ChunkPlan{};
}

template <class T> static int upload(T **dst, const std::vector<T> &src) {
  FH_HIP(hipMalloc(dst, sizeof(T) * std::max<size_t>(1, src.size())));
  if (!src.empty()) FH_HIP(hipMemcpy(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice));
  return FEMHOT_OK;
}

// Chunk plan for the LDS-staged scatter.  Elements are ordered by (segment, Morton code of the
// centroid) and cut into runs of at most `epc` elements that never cross a segment; inside a chunk
// the elements keep their original relative order (so per-element inputs/outputs are read and
// written in contiguous runs).  Each chunk lists its nodes by number of incident corners in the
// chunk (most first, ties by global id), flags the nodes shared with other chunks, stores its
// elements' chunk-local connectivity (uint16) and the coordinates of its nodes relative to its
// first node (FP32 and FP64).
// Modeled HBM bytes of one chunk-path call for a candidate element order (chunks are runs of at
// most epc elements within a segment; each run is sorted by element id here, as the plan keeps
// it): 128-byte lines of the per-element field read and of its gradient written through the
// order, lines of the nodal solution gathered per chunk, the per-chunk-node plan and output
// traffic (a shared node adds a partial-slot write and read per chunk), and the order itself
// unless it is the identity.
static double chunk_plan_cost(const Ctx *c, std::vector<int> &order, int epc, const std::vector<int> &seg_of,
                              bool identity) {
  const int n = c->n;
  const long long Ne = c->Ne;
  std::vector<int> loc, touch(c->Nn, 0);
  double bytes = identity ? 0.0 : 4.0 * (double)Ne;
  long long chunk_nodes = 0;
  for (long long i = 0; i < Ne;) {
    long long j = i;
    const int s = seg_of[order[i]];
    while (j < Ne && j - i < epc && seg_of[order[j]] == s) ++j;
    std::sort(order.begin() + i, order.begin() + j);
    long long last = -1, le = 0;
    for (long long k = i; k < j; ++k) {
      const long long l = (long long)order[k] * 4 / 128;
      if (l != last) ++le, last = l;
    }
    bytes += 2.0 * 128.0 * (double)le;
    loc.clear();
    for (long long k = i; k < j; ++k)
      for (int a = 0; a < n; ++a) loc.push_back(c->h_conn[(size_t)order[k] * n + a]);
    std::sort(loc.begin(), loc.end());
    loc.erase(std::unique(loc.begin(), loc.end()), loc.end());
    chunk_nodes += (long long)loc.size();
    last = -1;
    long long lu = 0;
    for (int g : loc) {
      touch[g]++;
      const long long l0 = (long long)g * 12 / 128, l1 = ((long long)g * 12 + 11) / 128;
      if (l0 > last) ++lu, last = l0;
      if (l1 > last) ++lu, last = l1;
    }
    bytes += 128.0 * (double)lu;
    i = j;
  }
  bytes += 22.0 * (double)chunk_nodes;
  for (long long g = 0; g < c->Nn; ++g)
    if (touch[g] == 1)
      bytes += 12.0;
    else if (touch[g] > 1)
      bytes += 24.0 * touch[g] + 12.0;
  return bytes;
}

// Host side of a chunk plan: the arrays build_chunk_plan uploads, and the figures the self-check
// (femhot_chunk_plan_check) reports.
struct HostChunkPlan {
  int epc = 0, n_chunks = 0, max_nodes = 0, max_elems = 0;
  bool perm_identity = false;
  std::vector<int> perm, elem_off, node_off, seg, nodes, zero, sh_nodes, node_slot, red_off, seg_chunk_off;
  std::vector<uint16_t> lconn, inc, rank, inc_start;
  std::vector<float> xyz32, xyz12;
  std::vector<double> xyz64;
};

static int plan_chunks_host(const Ctx *c, int epc, HostChunkPlan &h) {
  const int n = c->n, D = c->dim;
  const long long Ne = c->Ne;
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (long long i = 0; i < c->Nn; ++i)
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], c->h_coords[i * 3 + k]);
      hi[k] = std::max(hi[k], c->h_coords[i * 3 + k]);
    }
  std::vector<int> seg_of(Ne, 0);
  for (int s = 0; s < c->nseg; ++s)
    for (long long e = c->h_seg_off[s]; e < c->h_seg_off[s + 1]; ++e) seg_of[e] = s;
  std::vector<std::pair<uint64_t, int>> key(Ne);
  for (long long e = 0; e < Ne; ++e) {
    uint64_t q[3] = {0, 0, 0};
    for (int k = 0; k < D; ++k) {
      double s = 0;
      for (int a = 0; a < n; ++a) s += c->h_coords[(size_t)c->h_conn[e * n + a] * 3 + k];
      s /= n;
      const double span = hi[k] - lo[k];
      const double t = span > 0 ? (s - lo[k]) / span : 0.0;
      q[k] = (uint64_t)std::min(2097151.0, std::max(0.0, t * 2097151.0));
    }
    const uint64_t m = spread3(q[0]) | (spread3(q[1]) << 1) | (spread3(q[2]) << 2);
    key[e] = {m, (int)e};
  }
  std::stable_sort(key.begin(), key.end(), [&](const auto &x, const auto &y) {
    const int sx = seg_of[x.second], sy = seg_of[y.second];
    return sx != sy ? sx < sy : x.first < y.first;
  });
  h.epc = epc;
  h.perm.assign(Ne, 0);
  h.elem_off = {0};
  h.node_off = {0};
  h.lconn.assign((size_t)Ne * n, 0);
  h.inc.assign((size_t)Ne * n, 0);
  h.rank.assign((size_t)Ne * n, 0);
  std::vector<int> &perm = h.perm, &elem_off = h.elem_off, &node_off = h.node_off, &seg = h.seg, &nodes = h.nodes;
  std::vector<uint16_t> &lconn = h.lconn, &inc = h.inc, &rank = h.rank, &inc_start = h.inc_start;
  std::vector<int> cnt;
  std::vector<float> &xyz32 = h.xyz32, &xyz12 = h.xyz12;
  std::vector<double> &xyz64 = h.xyz64;
  // element order: spatial (Morton) chunks, or runs of the caller's own order when that order is
  // local enough to move fewer bytes (structured and mesher-ordered meshes often are)
  {
    std::vector<int> morton(Ne), ident(Ne);
    for (long long e = 0; e < Ne; ++e) morton[e] = key[e].second, ident[e] = (int)e;
    const double cost_m = chunk_plan_cost(c, morton, epc, seg_of, false);
    const double cost_i = chunk_plan_cost(c, ident, epc, seg_of, true);
    h.perm_identity = cost_i <= cost_m;
    perm = h.perm_identity ? ident : morton;
  }
  long long i = 0;
  std::vector<int> loc, loc2, lid, icount, lorder, lrank;
  int max_nodes = 0, max_elems = 0;
  while (i < Ne) {
    long long j = i;
    const int s = seg_of[perm[i]];
    while (j < Ne && j - i < epc && seg_of[perm[j]] == s) ++j;
    std::sort(perm.begin() + i, perm.begin() + j);
    loc.clear();
    for (long long k = i; k < j; ++k)
      for (int a = 0; a < n; ++a) loc.push_back(c->h_conn[(size_t)perm[k] * n + a]);
    std::sort(loc.begin(), loc.end());
    loc.erase(std::unique(loc.begin(), loc.end()), loc.end());
    if (loc.size() > 65535) {
      set_error("ensure_chunks: more than 65535 nodes in a chunk");
      return FEMHOT_ERR_UNSUPPORTED;
    }
    if ((j - i) * n > 65535) {
      set_error("ensure_chunks: too many element corners in a chunk");
      return FEMHOT_ERR_UNSUPPORTED;
    }
    // Local node order: by number of incident corners in the chunk, most first (ties by node id).
    // The node gather gives one lane per node and walks its incidence list, so a wave then walks
    // lists of similar length instead of waiting on its longest one.
    lid.resize((size_t)(j - i) * n);
    icount.assign(loc.size(), 0);
    for (long long k = i; k < j; ++k)
      for (int a = 0; a < n; ++a) {
        const int g = c->h_conn[(size_t)perm[k] * n + a];
        const int q = (int)(std::lower_bound(loc.begin(), loc.end(), g) - loc.begin());
        lid[(size_t)(k - i) * n + a] = q;
        icount[q]++;
      }
    lorder.resize(loc.size());
    std::iota(lorder.begin(), lorder.end(), 0);
    std::stable_sort(lorder.begin(), lorder.end(), [&](int x, int y) { return icount[x] > icount[y]; });
    lrank.resize(loc.size());
    loc2.resize(loc.size());
    for (size_t q = 0; q < loc.size(); ++q) {
      lrank[lorder[q]] = (int)q;
      loc2[q] = loc[lorder[q]];
    }
    loc.swap(loc2);
    for (long long k = i; k < j; ++k)
      for (int a = 0; a < n; ++a) lconn[(size_t)k * n + a] = (uint16_t)lrank[lid[(size_t)(k - i) * n + a]];
    // incidence (counting sort by local node; entries of a node in element-position order)
    cnt.assign(loc.size() + 1, 0);
    for (long long k = i; k < j; ++k)
      for (int a = 0; a < n; ++a) cnt[lconn[(size_t)k * n + a] + 1]++;
    for (size_t q = 0; q < loc.size(); ++q) cnt[q + 1] += cnt[q];
    for (size_t q = 0; q < loc.size(); ++q) inc_start.push_back((uint16_t)cnt[q]);
    for (long long k = i; k < j; ++k)
      for (int a = 0; a < n; ++a) {
        const int q = lconn[(size_t)k * n + a];
        rank[(size_t)k * n + a] = (uint16_t)cnt[q];  // inc[i * n + rank] is this corner
        inc[(size_t)i * n + cnt[q]++] = (uint16_t)((k - i) * n + a);
      }
    const double *ref = &c->h_coords[(size_t)loc[0] * 3];
    for (int g : loc)
      for (int k = 0; k < 4; ++k) {  // (x, y, z) per node; the 16-byte FP32 record adds a 0
        const double v = k < 3 ? c->h_coords[(size_t)g * 3 + k] - ref[k] : 0.0;
        xyz32.push_back((float)v);
        if (k < 3) xyz12.push_back((float)v), xyz64.push_back(v);
      }
    nodes.insert(nodes.end(), loc.begin(), loc.end());
    max_nodes = std::max<int>(max_nodes, (int)loc.size());
    max_elems = std::max<int>(max_elems, (int)(j - i));
    elem_off.push_back((int)j);
    node_off.push_back((int)nodes.size());
    seg.push_back(s);
    i = j;
  }
  std::vector<int> touch(c->Nn, 0);
  for (int g : nodes) touch[g]++;
  std::vector<int> &zero = h.zero, &sh_nodes = h.sh_nodes;
  std::vector<int> sh_index(c->Nn, -1);
  for (long long g = 0; g < c->Nn; ++g) {
    if (touch[g] == 0) zero.push_back((int)g);
    if (touch[g] > 1) {
      sh_index[g] = (int)sh_nodes.size();
      sh_nodes.push_back((int)g);
    }
  }
  // partial slots: shared node k (in chunks c_0 < c_1 < ...) owns slot k for c_0, n_shared + k for c_1
  // and 2 n_shared + red_off[k] - 2 k + o - 2 for c_o, o >= 2 (red_off: prefix sums of the chunk
  // counts), so the reduction finds every shared node's first two partials without a table
  h.node_slot.assign(nodes.size(), -1);
  h.red_off.assign(sh_nodes.size() + 1, 0);
  std::vector<int> &node_slot = h.node_slot, &red_off = h.red_off;
  for (int g : sh_nodes) red_off[sh_index[g] + 1] = touch[g];
  for (size_t k = 0; k < sh_nodes.size(); ++k) red_off[k + 1] += red_off[k];
  {
    const int nsh = (int)sh_nodes.size();
    std::vector<int> occ(sh_nodes.size(), 0);
    for (size_t k = 0; k < nodes.size(); ++k) {  // nodes are listed chunk by chunk, in chunk order
      const int g = nodes[k];
      if (touch[g] < 2) continue;
      const int q = sh_index[g], o = occ[q]++;
      node_slot[k] = o == 0 ? q : o == 1 ? nsh + q : 2 * nsh + red_off[q] - 2 * q + o - 2;
    }
  }
  h.seg_chunk_off.assign(c->nseg + 1, 0);
  std::vector<int> &seg_chunk_off = h.seg_chunk_off;
  for (int s : seg) seg_chunk_off[s + 1]++;
  for (int s = 0; s < c->nseg; ++s) seg_chunk_off[s + 1] += seg_chunk_off[s];
  h.n_chunks = (int)seg.size();
  h.max_nodes = max_nodes;
  h.max_elems = max_elems;
  // the pipelined kernel reads inc_start as whole 32-bit words: one entry of padding past the end
  inc_start.push_back(0);
  inc_start.push_back(0);
  return FEMHOT_OK;
}

static int build_chunk_plan(Ctx *c, int epc, ChunkPlan &p) {
  HostChunkPlan h;
  if (int rc = plan_chunks_host(c, epc, h)) return rc;
  p.epc = epc;
  p.n_chunks = h.n_chunks;
  p.max_nodes = h.max_nodes;
  p.max_elems = h.max_elems;
  p.perm_identity = h.perm_identity;
  p.n_shared = (long long)h.sh_nodes.size();
  p.n_zero = (long long)h.zero.size();
  p.n_slots = (long long)h.red_off.back();
  int rc;
  if ((rc = upload(&p.chunk_elem_off, h.elem_off)) || (rc = upload(&p.chunk_node_off, h.node_off)) ||
      (rc = upload(&p.chunk_seg, h.seg)) || (rc = upload(&p.perm, h.perm)) || (rc = upload(&p.lconn, h.lconn)) ||
      (rc = upload(&p.chunk_nodes, h.nodes)) || (rc = upload(&p.chunk_xyz32, h.xyz32)) ||
      (rc = upload(&p.chunk_xyz12, h.xyz12)) || (rc = upload(&p.chunk_xyz64, h.xyz64)) ||
      (rc = upload(&p.zero_nodes, h.zero)) || (rc = upload(&p.node_slot, h.node_slot)) ||
      (rc = upload(&p.sh_nodes, h.sh_nodes)) || (rc = upload(&p.red_off, h.red_off)) ||
      (rc = upload(&p.seg_chunk_off, h.seg_chunk_off)) || (rc = upload(&p.inc, h.inc)) ||
      (rc = upload(&p.inc_start, h.inc_start)) || (rc = upload(&p.rank, h.rank))) {
    free_chunks(p);
    return rc;
  }
  return FEMHOT_OK;
}

// Every invariant the chunk kernels rely on, checked on a host plan; "" when all hold, else the first
// violation.  The kernels index LDS and the partial slots with these arrays unchecked.
static std::string check_chunk_plan(const Ctx &c, const HostChunkPlan &h) {
  const int n = c.n;
  const long long Ne = c.Ne, Nn = c.Nn;
  auto at = [](const char *what, long long i) { return std::string(what) + " (at " + std::to_string(i) + ")"; };
  const int K = h.n_chunks;
  if ((int)h.elem_off.size() != K + 1 || (int)h.node_off.size() != K + 1 || (int)h.seg.size() != K)
    return "chunk offset arrays do not have n_chunks + 1 entries";
  if (h.elem_off[0] != 0 || h.elem_off[K] != Ne) return "element offsets do not span [0, n_elems)";
  if (h.node_off[0] != 0 || (size_t)h.node_off[K] != h.nodes.size()) return "node offsets do not span the node list";
  if ((long long)h.perm.size() != Ne || (long long)h.lconn.size() != Ne * n || (long long)h.rank.size() != Ne * n ||
      (long long)h.inc.size() != Ne * n)
    return "per-element arrays have the wrong length";
  if (h.inc_start.size() != h.nodes.size() + 2) return "incidence starts: one entry per chunk node plus 2 of padding";
  if (h.xyz64.size() != 3 * h.nodes.size() || h.xyz12.size() != 3 * h.nodes.size() ||
      h.xyz32.size() != 4 * h.nodes.size())
    return "coordinate records have the wrong length";
  std::vector<char> seen(Ne, 0);
  for (long long e = 0; e < Ne; ++e) {
    const int q = h.perm[e];
    if (q < 0 || q >= Ne || seen[q]) return at("perm is not a permutation", e);
    if (h.perm_identity && q != e) return at("perm_identity set but perm[i] != i", e);
    seen[q] = 1;
  }
  std::vector<int> touch(Nn, 0), last_chunk(Nn, -1), icount, hit;
  int max_nodes = 0, max_elems = 0;
  for (int k = 0; k < K; ++k) {
    const int e0 = h.elem_off[k], e1 = h.elem_off[k + 1], n0 = h.node_off[k], n1 = h.node_off[k + 1];
    const int ne = e1 - e0, nl = n1 - n0, s = h.seg[k];
    if (ne <= 0 || ne > h.epc) return at("chunk size outside [1, epc]", k);
    if (nl <= 0) return at("chunk without nodes", k);
    if (s < 0 || s >= c.nseg || (k > 0 && s < h.seg[k - 1])) return at("chunk segment out of order", k);
    max_nodes = std::max(max_nodes, nl);
    max_elems = std::max(max_elems, ne);
    for (int i = n0; i < n1; ++i) {
      const int g = h.nodes[i];
      if (g < 0 || g >= Nn) return at("chunk node id out of range", i);
      if (last_chunk[g] == k) return at("node listed twice in one chunk", i);
      last_chunk[g] = k;
      touch[g]++;
      const double *ref = &c.h_coords[(size_t)h.nodes[n0] * 3];
      for (int d = 0; d < 3; ++d) {
        const double v = c.h_coords[(size_t)g * 3 + d] - ref[d];
        if (h.xyz64[(size_t)i * 3 + d] != v || h.xyz12[(size_t)i * 3 + d] != (float)v ||
            h.xyz32[(size_t)i * 4 + d] != (float)v)
          return at("coordinate record differs from the chunk-relative coordinates", i);
      }
      if (h.xyz32[(size_t)i * 4 + 3] != 0.0f) return at("16-byte coordinate record not padded with 0", i);
    }
    icount.assign(nl, 0);
    for (int e = e0; e < e1; ++e) {
      const long long ge = h.perm[e];
      if (ge < c.h_seg_off[s] || ge >= c.h_seg_off[s + 1]) return at("element outside its chunk's segment", e);
      for (int a = 0; a < n; ++a) {
        const int l = h.lconn[(size_t)e * n + a];
        if (l >= nl) return at("local node index beyond the chunk's nodes", e);
        if (h.nodes[n0 + l] != c.h_conn[ge * n + a]) return at("local connectivity names another node", e);
        icount[l]++;
      }
    }
    for (int q = 0; q < nl; ++q) {
      if (icount[q] == 0) return at("chunk node in none of the chunk's elements", n0 + q);
      if (q > 0 && icount[q] > icount[q - 1]) return at("chunk nodes not ordered by incident corners", n0 + q);
      const int js = h.inc_start[n0 + q], je = q + 1 < nl ? h.inc_start[n0 + q + 1] : n * ne;
      if (js != (q == 0 ? 0 : (int)h.inc_start[n0 + q - 1] + icount[q - 1]) || je - js != icount[q])
        return at("incidence run does not match the node's corner count", n0 + q);
    }
    hit.assign((size_t)n * ne, 0);
    for (int e = e0; e < e1; ++e)
      for (int a = 0; a < n; ++a) {
        const int l = h.lconn[(size_t)e * n + a], r = h.rank[(size_t)e * n + a];
        const int js = h.inc_start[n0 + l], je = l + 1 < nl ? h.inc_start[n0 + l + 1] : n * ne;
        if (r < js || r >= je) return at("corner rank outside its node's incidence run", e);
        if (hit[r]++) return at("two corners share one incidence position", e);
        if (h.inc[(size_t)e0 * n + r] != (e - e0) * n + a) return at("incidence entry does not name its corner", e);
        if (r > js && h.inc[(size_t)e0 * n + r - 1] >= h.inc[(size_t)e0 * n + r])
          return at("incidence run not in element-position order", e);
      }
  }
  if (max_nodes != h.max_nodes || max_elems != h.max_elems) return "max_nodes / max_elems differ from the chunks";
  for (int s = 0; s < c.nseg; ++s)
    for (int k = h.seg_chunk_off[s]; k < h.seg_chunk_off[s + 1]; ++k)
      if (h.seg[k] != s) return at("seg_chunk_off names a chunk of another segment", k);
  if (h.seg_chunk_off[c.nseg] != K) return "seg_chunk_off does not end at n_chunks";
  // zero / shared nodes, partial slots: a shared node owns one slot per chunk containing it, in
  // chunk order, contiguous; exclusive nodes own none
  size_t iz = 0, ish = 0;
  for (long long g = 0; g < Nn; ++g) {
    if (touch[g] == 0 && (iz >= h.zero.size() || h.zero[iz++] != g)) return at("zero_nodes misses a node", g);
    if (touch[g] > 1) {
      if (ish >= h.sh_nodes.size() || h.sh_nodes[ish] != g) return at("sh_nodes misses a shared node", g);
      if (h.red_off[ish + 1] - h.red_off[ish] != touch[g]) return at("shared node's slot count", g);
      ++ish;
    }
  }
  if (iz != h.zero.size() || ish != h.sh_nodes.size()) return "zero_nodes / sh_nodes list extra nodes";
  const long long nsh = (long long)h.sh_nodes.size();
  std::vector<int> sh_index(Nn, -1), occ(h.sh_nodes.size(), 0);
  std::vector<char> used(h.red_off.back(), 0);
  for (size_t k = 0; k < h.sh_nodes.size(); ++k) sh_index[h.sh_nodes[k]] = (int)k;
  for (int k = 0; k < K; ++k)
    for (int i = h.node_off[k]; i < h.node_off[k + 1]; ++i) {
      const int g = h.nodes[i], sl = h.node_slot[i];
      if (touch[g] == 1) {
        if (sl != -1) return at("exclusive node owns a partial slot", i);
        continue;
      }
      const long long q = sh_index[g], o = occ[q]++;
      const long long want = o == 0 ? q : o == 1 ? nsh + q : 2 * nsh + h.red_off[q] - 2 * q + o - 2;
      if (sl != want) return at("partial slot of a shared node not at its chunk-order place", i);
      if (sl < 0 || sl >= (long long)used.size() || used[sl]++) return at("partial slot out of range or reused", i);
    }
  return "";
}

// Plans are cached per elements-per-chunk, so call shapes tuned to different chunk sizes (a forward
// and its VJP, FP32 and FP64 calls) alternate without rebuilding.  A plan is built on the host and
// uploaded with blocking copies, so it is never built while `s` is being captured into a graph.
int ensure_chunks(Ctx *c, int epc, hipStream_t s, const ChunkPlan **out) {
  for (ChunkPlan &p : c->plans)
    if (p.epc == epc && p.n_chunks) {
      p.last_use = ++c->plan_tick;
      if (out) *out = &p;
      return FEMHOT_OK;
    }
  if (epc < 1 || epc > 8192) {
    set_error("ensure_chunks: elements per chunk must be in [1, 8192]");
    return FEMHOT_ERR_ARG;
  }
  if (s) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    FH_HIP(hipStreamIsCapturing(s, &cs));
    if (cs != hipStreamCaptureStatusNone) {
      set_error("ensure_chunks: the chunk plan of " + std::to_string(epc) +
                " elements per chunk is not built; build it before capture with an uncaptured call of the same "
                "shape or femhot_set_chunk_config");
      return FEMHOT_ERR_ARG;
    }
  }
  if (c->plans.size() >= kMaxChunkPlans) {  // evict the least recently used plan
    auto victim = c->plans.begin();
    for (auto it = c->plans.begin(); it != c->plans.end(); ++it)
      if (it->last_use < victim->last_use) victim = it;
    FH_HIP(hipDeviceSynchronize());
    free_chunks(*victim);
    c->plans.erase(victim);
  }
  c->plans.emplace_back();
  ChunkPlan &p = c->plans.back();
  const int rc = build_chunk_plan(c, epc, p);
  if (rc) {
    c->plans.pop_back();
    return rc;
  }
  p.last_use = ++c->plan_tick;
  if (out) *out = &p;
  return FEMHOT_OK;
}

const ChunkPlan *find_chunks(Ctx *c, int epc) {
  for (ChunkPlan &p : c->plans)
    if (p.epc == epc && p.n_chunks) {
      p.last_use = ++c->plan_tick;
      return &p;
    }
  return nullptr;
}

void free_chunk_plans(Ctx *c) {
  for (ChunkPlan &p : c->plans) free_chunks(p);
  c->plans.clear();
}

int timing_mark(Ctx *c, hipStream_t s) {
  if (!c->timing) return FEMHOT_OK;
  if (c->tev_used == c->tev_pool.size()) {
    hipEvent_t e;
    FH_HIP(hipEventCreate(&e));
    c->tev_pool.push_back(e);
  }
  FH_HIP(hipEventRecord(c->tev_pool[c->tev_used++], s));
  return FEMHOT_OK;
}

int ensure_workspace(Ctx *c, size_t bytes) {
  if (c->ws_bytes >= bytes) return FEMHOT_OK;
  if (c->d_ws) FH_HIP(hipFree(c->d_ws));
  c->d_ws = nullptr;
  c->ws_bytes = 0;
  FH_HIP(hipMalloc(&c->d_ws, bytes));
  c->ws_bytes = bytes;
  return FEMHOT_OK;
}

static double det_at_center(int et, const double *X) {
  // J at the one-point rule (mesh.py:128-132 uses GetIntegrationData of the default rule)
  const int n = nodes_per_elem(et), D = elem_dim(et);
  double J[3][3] = {{0}};
  double gx, gy, gz, w;
  if (et == TET4)
    gauss_point<TET4, 1>(0, gx, gy, gz, w);
  else
    gx = gy = gz = 0.0;
  for (int a = 0; a < n; ++a) {
    double d[3];
    if (et == QUAD4) {
      const int sx[4] = {-1, 1, 1, -1}, sy[4] = {-1, -1, 1, 1};
      d[0] = 0.25 * sx[a] * (1 + sy[a] * gy);
      d[1] = 0.25 * sy[a] * (1 + sx[a] * gx);
    } else if (et == HEX8) {
      const int sx[8] = {-1, 1, 1, -1, -1, 1, 1, -1}, sy[8] = {-1, -1, 1, 1, -1, -1, 1, 1},
                sz[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
      d[0] = 0.125 * sx[a] * (1 + sy[a] * gy) * (1 + sz[a] * gz);
      d[1] = 0.125 * sy[a] * (1 + sx[a] * gx) * (1 + sz[a] * gz);
      d[2] = 0.125 * sz[a] * (1 + sx[a] * gx) * (1 + sy[a] * gy);
    } else {
      const double t[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
      d[0] = t[a][0];
      d[1] = t[a][1];
      d[2] = t[a][2];
    }
    for (int i = 0; i < D; ++i)
      for (int k = 0; k < D; ++k) J[i][k] += X[a * 3 + i] * d[k];
  }
  if (D == 2) return J[0][0] * J[1][1] - J[0][1] * J[1][0];
  return J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
         J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
}

}  // namespace femhot

using namespace femhot;

// Host-side part of a context: the checks on the mesh description, the host copies of the mesh and
// the physics constants (no device memory).  On failure the error is set.
static int init_host_ctx(const femhot_mesh_desc *d, Ctx &c) {
  FH_CHECK_ARG(d->elem_type >= QUAD4 && d->elem_type <= TET4, FEMHOT_ERR_ARG, "unknown element type");
  FH_CHECK_ARG(d->physics >= THERMAL && d->physics <= NEOHOOKE, FEMHOT_ERR_ARG, "unknown physics");
  FH_CHECK_ARG(d->num_gp >= 0 && d->num_gp <= 3, FEMHOT_ERR_ARG, "num_gp must be 0 (default), 1, 2 or 3");
  FH_CHECK_ARG(d->n_nodes > 0 && d->n_elems > 0, FEMHOT_ERR_SHAPE, "empty mesh");
  FH_CHECK_ARG(d->n_nodes < (1LL << 31) && d->n_elems < (1LL << 31), FEMHOT_ERR_SHAPE, "mesh too large for int32");
  FH_CHECK_ARG(d->coords && d->conn, FEMHOT_ERR_ARG, "null coords/conn");
  FH_CHECK_ARG(d->n_dirichlet == 0 || (d->dirichlet_idx && d->dirichlet_val), FEMHOT_ERR_ARG,
               "null Dirichlet arrays");
  c.desc = *d;
  c.et = d->elem_type;
  c.ph = d->physics;
  c.ngp = d->num_gp ? d->num_gp : default_num_gp(c.et);
  c.n = nodes_per_elem(c.et);
  c.dim = elem_dim(c.et);
  c.ndpn = c.ph == THERMAL ? 1 : c.dim;
  c.Nn = d->n_nodes;
  c.Ne = d->n_elems;
  c.Ndof = c.Nn * c.ndpn;
  c.h_coords.assign(d->coords, d->coords + c.Nn * 3);
  c.h_conn.assign(d->conn, d->conn + c.Ne * c.n);
  for (long long i = 0; i < c.Ne * c.n; ++i)
    if (c.h_conn[i] < 0 || c.h_conn[i] >= c.Nn) {
      set_error("femhot_ctx_create: connectivity index out of range");
      return FEMHOT_ERR_INDEX;
    }
  // collapsed elements (a node listed twice) are valid input, but two lanes of one element would
  // then update the same tangent entries: the colour-ordered tangent falls back to atomics for them
  for (long long e = 0; e < c.Ne && !c.repeated_nodes; ++e)
    for (int a = 0; a < c.n && !c.repeated_nodes; ++a)
      for (int b = a + 1; b < c.n; ++b)
        if (c.h_conn[e * c.n + a] == c.h_conn[e * c.n + b]) {
          c.repeated_nodes = true;
          break;
        }
  for (long long i = 0; i < c.Nn * 3; ++i)
    if (!std::isfinite(c.h_coords[i])) {
      set_error("femhot_ctx_create: non-finite coordinate");
      return FEMHOT_ERR_NONFINITE;
    }
  c.h_dir_idx.assign(d->dirichlet_idx, d->dirichlet_idx + d->n_dirichlet);
  c.h_dir_val.assign(d->dirichlet_val, d->dirichlet_val + d->n_dirichlet);
  for (long long v : c.h_dir_idx)
    if (v < 0 || v >= c.Ndof) {
      set_error("femhot_ctx_create: Dirichlet dof index out of range");
      return FEMHOT_ERR_INDEX;
    }
  c.nseg = d->n_segments > 1 ? d->n_segments : 1;
  if (c.nseg > 1) {
    if (!d->segment_offsets) {
      set_error("femhot_ctx_create: segment_offsets required when n_segments > 1");
      return FEMHOT_ERR_ARG;
    }
    c.h_seg_off.assign(d->segment_offsets, d->segment_offsets + c.nseg + 1);
    bool ok = c.h_seg_off[0] == 0 && c.h_seg_off[c.nseg] == c.Ne;
    for (int s = 0; s < c.nseg; ++s) ok = ok && c.h_seg_off[s] <= c.h_seg_off[s + 1];
    if (!ok) {
      set_error("femhot_ctx_create: segment_offsets must be non-decreasing from 0 to n_elems");
      return FEMHOT_ERR_SHAPE;
    }
  } else {
    c.h_seg_off = {0, c.Ne};
  }
  // physics parameters (mechanical.py:60-82, mechanical_neohooke.py:253-255)
  PhysParams &p = c.pp;
  p.E = d->young_modulus;
  p.nu = d->poisson_ratio;
  p.beta = d->beta;
  p.c = d->c;
  for (int k = 0; k < 3; ++k) p.body[k] = d->body_force[k];
  p.has_body = (p.body[0] != 0.0 || p.body[1] != 0.0 || p.body[2] != 0.0) ? 1 : 0;
  const double E = p.E, nu = p.nu;
  const double c1 = E / ((1.0 + nu) * (1.0 - 2.0 * nu));
  p.c2 = c1 * (1.0 - nu);
  p.c3 = c1 * nu;
  p.c4 = c1 * 0.5 * (1.0 - 2.0 * nu);
  p.f2 = E / (1.0 - nu * nu);
  p.f2nu = p.f2 * nu;
  p.f2sh = p.f2 * (1.0 - nu) / 2.0;
  p.kap1 = 1.0 / (3.0 * (1.0 - 2.0 * nu));
  p.mu1 = 1.0 / (2.0 * (1.0 + nu));
  return FEMHOT_OK;
}

extern "C" {

const char *femhot_last_error(void) { return get_error().c_str(); }
const char *femhot_version(void) { return "femhot 0.1 gfx950"; }

int femhot_ctx_create(const femhot_mesh_desc *d, femhot_ctx **out) {
  FH_CHECK_ARG(d && out, FEMHOT_ERR_ARG, "null argument");
  *out = nullptr;
  auto *h = new femhot_ctx();
  Ctx &c = h->c;
  if (int rc = init_host_ctx(d, c)) {
    delete h;
    return rc;
  }
  // device copies
  std::vector<float> c32(c.h_coords.begin(), c.h_coords.end());
  std::vector<uint8_t> dmask(c.Ndof, 0);
  for (long long v : c.h_dir_idx) dmask[v] = 1;
  auto fail = [&](hipError_t e, const char *what) {
    set_error(std::string("femhot_ctx_create: ") + what + " -> " + hipGetErrorString(e));
    femhot_ctx_destroy(h);
    return FEMHOT_ERR_HIP;
  };
  hipError_t e;
  if ((e = hipMalloc(&c.d_coords64, sizeof(double) * c.Nn * 3)) != hipSuccess) return fail(e, "hipMalloc");
  if ((e = hipMalloc(&c.d_coords32, sizeof(float) * c.Nn * 3)) != hipSuccess) return fail(e, "hipMalloc");
  if ((e = hipMalloc(&c.d_conn, sizeof(int) * c.Ne * c.n)) != hipSuccess) return fail(e, "hipMalloc");
  if ((e = hipMalloc(&c.d_dmask, c.Ndof)) != hipSuccess) return fail(e, "hipMalloc");
  if ((e = hipMalloc(&c.d_seg_off, sizeof(long long) * (c.nseg + 1))) != hipSuccess) return fail(e, "hipMalloc");
  if ((e = hipMemcpy(c.d_coords64, c.h_coords.data(), sizeof(double) * c.Nn * 3, hipMemcpyHostToDevice)) != hipSuccess)
    return fail(e, "hipMemcpy");
  if ((e = hipMemcpy(c.d_coords32, c32.data(), sizeof(float) * c.Nn * 3, hipMemcpyHostToDevice)) != hipSuccess)
    return fail(e, "hipMemcpy");
  if ((e = hipMemcpy(c.d_conn, c.h_conn.data(), sizeof(int) * c.Ne * c.n, hipMemcpyHostToDevice)) != hipSuccess)
    return fail(e, "hipMemcpy");
  if ((e = hipMemcpy(c.d_dmask, dmask.data(), c.Ndof, hipMemcpyHostToDevice)) != hipSuccess) return fail(e, "hipMemcpy");
  if ((e = hipMemcpy(c.d_seg_off, c.h_seg_off.data(), sizeof(long long) * (c.nseg + 1), hipMemcpyHostToDevice)) !=
      hipSuccess)
    return fail(e, "hipMemcpy");
  if ((e = hipHostMalloc((void **)&c.h_pinned, 1024 * sizeof(double), hipHostMallocDefault)) != hipSuccess)
    return fail(e, "hipHostMalloc");
  if (int rc = init_fixed_scratch()) {  // this device's address of the allocation-free path's scratch
    femhot_ctx_destroy(h);
    return rc;
  }
  *out = h;
  return FEMHOT_OK;
}

void femhot_ctx_destroy(femhot_ctx *h) {
  if (!h) return;
  forget_tuning(&h->c);
  Ctx &c = h->c;
  (void)hipFree(c.d_coords64);
  (void)hipFree(c.d_coords32);
  (void)hipFree(c.d_conn);
  (void)hipFree(c.d_dmask);
  (void)hipFree(c.d_seg_off);
  (void)hipFree(c.d_row_ptr);
  (void)hipFree(c.d_col_idx);
  (void)hipFree(c.d_slot);
  (void)hipFree(c.d_ne_off);
  (void)hipFree(c.d_diag_off);
  (void)hipFree(c.d_ne_list);
  (void)hipFree(c.d_colour_elems);
  free_chunk_plans(&c);
  (void)hipFree(c.chunk_ws.ptr);
  (void)hipFree(c.newton_ws.ptr);
  (void)hipFree(c.d_ws);
  for (hipEvent_t e : c.tev_pool) (void)hipEventDestroy(e);
  if (c.h_pinned) (void)hipHostFree(c.h_pinned);
  delete h;
}

int femhot_ctx_info(const femhot_ctx *h, long long *n_dofs, int *npe, int *dpn, long long *nnz, int *n_colours,
                    int *n_segments) {
  FH_CHECK_ARG(h, FEMHOT_ERR_ARG, "null context");
  Ctx *c = const_cast<Ctx *>(&h->c);
  if (n_dofs) *n_dofs = c->Ndof;
  if (npe) *npe = c->n;
  if (dpn) *dpn = c->ndpn;
  if (nnz) {
    int rc = ensure_pattern(c);
    if (rc) return rc;
    *nnz = c->nnz;
  }
  if (n_colours) {
    int rc = ensure_colouring(c);
    if (rc) return rc;
    *n_colours = c->n_colours;
  }
  if (n_segments) *n_segments = c->nseg;
  return FEMHOT_OK;
}

int femhot_pattern(femhot_ctx *h, long long *row_ptr, int *col_idx, void *stream) {
  FH_CHECK_ARG(h && row_ptr && col_idx, FEMHOT_ERR_ARG, "null argument");
  Ctx *c = &h->c;
  int rc = ensure_pattern(c);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  FH_HIP(hipMemcpyAsync(row_ptr, c->d_row_ptr, sizeof(long long) * (c->Ndof + 1), hipMemcpyDeviceToDevice, s));
  if (c->nnz) FH_HIP(hipMemcpyAsync(col_idx, c->d_col_idx, sizeof(int) * c->nnz, hipMemcpyDeviceToDevice, s));
  return FEMHOT_OK;
}

int femhot_colouring(femhot_ctx *h, int *colour) {
  FH_CHECK_ARG(h && colour, FEMHOT_ERR_ARG, "null argument");
  Ctx *c = &h->c;
  int rc = ensure_colouring(c);
  if (rc) return rc;
  std::memcpy(colour, c->h_colour.data(), sizeof(int) * c->Ne);
  return FEMHOT_OK;
}

int femhot_kernel_timing(femhot_ctx *h, int enable, double *total_ms, long long *launches) {
  FH_CHECK_ARG(h, FEMHOT_ERR_ARG, "null context");
  Ctx &c = h->c;
  if (enable) {
    c.timing = true;
    c.tev_used = 0;
    return FEMHOT_OK;
  }
  double ms = 0.0;
  long long n = 0;
  for (size_t i = 0; i + 1 < c.tev_used; i += 2) {
    FH_HIP(hipEventSynchronize(c.tev_pool[i + 1]));
    float t = 0.f;
    FH_HIP(hipEventElapsedTime(&t, c.tev_pool[i], c.tev_pool[i + 1]));
    ms += t;
    ++n;
  }
  c.timing = false;
  c.tev_used = 0;
  if (total_ms) *total_ms = ms;
  if (launches) *launches = n;
  return FEMHOT_OK;
}

int femhot_chunk_plan_check(const femhot_mesh_desc *d, int epc, long long *stats) {
  FH_CHECK_ARG(d && stats, FEMHOT_ERR_ARG, "null argument");
  FH_CHECK_ARG(epc > 0, FEMHOT_ERR_ARG, "elements per chunk must be positive");
  Ctx c;
  if (int rc = init_host_ctx(d, c)) return rc;
  HostChunkPlan h;
  if (int rc = plan_chunks_host(&c, epc, h)) return rc;
  const std::string why = check_chunk_plan(c, h);
  const long long v[8] = {h.n_chunks, h.max_nodes, h.max_elems, (long long)h.nodes.size(),
                          (long long)h.sh_nodes.size(), (long long)h.red_off.back(), (long long)h.zero.size(),
                          h.perm_identity ? 1 : 0};
  for (int k = 0; k < 8; ++k) stats[k] = v[k];
  if (!why.empty()) {
    set_error("femhot_chunk_plan_check: " + why);
    return FEMHOT_ERR_INDEX;
  }
  return FEMHOT_OK;
}

int femhot_orient_elements(int et, long long n_nodes, const double *coords, long long n_elems, int *conn,
                           long long *n_swapped) {
  FH_CHECK_ARG(et >= QUAD4 && et <= TET4, FEMHOT_ERR_ARG, "unknown element type");
  FH_CHECK_ARG(coords && conn, FEMHOT_ERR_ARG, "null argument");
  const int n = nodes_per_elem(et);
  long long sw = 0;
  double X[8 * 3];
  for (long long e = 0; e < n_elems; ++e) {
    for (int a = 0; a < n; ++a) {
      const int g = conn[e * n + a];
      FH_CHECK_ARG(g >= 0 && g < n_nodes, FEMHOT_ERR_INDEX, "connectivity index out of range");
      for (int k = 0; k < 3; ++k) X[a * 3 + k] = coords[(size_t)g * 3 + k];
    }
    if (det_at_center(et, X) < 0.0) {
      std::swap(conn[e * n + 0], conn[e * n + 1]);
      ++sw;
    }
  }
  if (n_swapped) *n_swapped = sw;
  return FEMHOT_OK;
}

}  // extern "C"
