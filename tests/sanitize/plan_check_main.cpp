// Host sanitizer run of the chunk planner: femhot_chunk_plan_check (host only, no GPU) on structured
// and shuffled QUAD4 / HEX8 / TET4 meshes, uneven segments with an empty one, unused nodes and a
// collapsed element, at chunk sizes from 1 element up, plus femhot_orient_elements.  Built with
// AddressSanitizer and UndefinedBehaviorSanitizer by the Makefile beside it; exits non-zero on the
// first failed check (a sanitizer report aborts the run by itself).
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "femhot.h"

namespace {

struct Mesh {
  int et;
  std::vector<double> xyz;  // 3 per node
  std::vector<int> conn;
  long long nn() const { return (long long)xyz.size() / 3; }
  int npe() const { return et == FEMHOT_QUAD4 ? 4 : et == FEMHOT_HEX8 ? 8 : 4; }
  long long ne() const { return (long long)conn.size() / npe(); }
};

// (nx+1)(ny+1)(nz+1) grid nodes with a seeded jitter; nz = 0 gives a plane
std::vector<double> grid(int nx, int ny, int nz, double jitter, unsigned seed) {
  std::mt19937 rng(seed);
  std::uniform_real_distribution<double> u(-jitter, jitter);
  std::vector<double> x;
  for (int k = 0; k <= nz; ++k)
    for (int j = 0; j <= ny; ++j)
      for (int i = 0; i <= nx; ++i) {
        x.push_back(i + u(rng));
        x.push_back(j + u(rng));
        x.push_back(nz ? k + u(rng) : 0.0);
      }
  return x;
}

Mesh quads(int n) {
  Mesh m{FEMHOT_QUAD4, grid(n, n, 0, 0.2, 1u), {}};
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      const int a = j * (n + 1) + i;
      for (int v : {a, a + 1, a + n + 2, a + n + 1}) m.conn.push_back(v);
    }
  return m;
}

int node3(int i, int j, int k, int nx, int ny) { return (k * (ny + 1) + j) * (nx + 1) + i; }

Mesh hexes(int nx, int ny, int nz) {
  Mesh m{FEMHOT_HEX8, grid(nx, ny, nz, 0.15, 2u), {}};
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const int c[8] = {node3(i, j, k, nx, ny),         node3(i + 1, j, k, nx, ny),
                          node3(i + 1, j + 1, k, nx, ny), node3(i, j + 1, k, nx, ny),
                          node3(i, j, k + 1, nx, ny),     node3(i + 1, j, k + 1, nx, ny),
                          node3(i + 1, j + 1, k + 1, nx, ny), node3(i, j + 1, k + 1, nx, ny)};
        m.conn.insert(m.conn.end(), c, c + 8);
      }
  return m;
}

// each hex of an n^3 grid split into the 6 tetrahedra around its main diagonal (corner 0 to 6)
Mesh tets(int n, unsigned seed) {
  Mesh m{FEMHOT_TET4, grid(n, n, n, 0.15, seed), {}};
  const int paths[6][2] = {{1, 2}, {1, 5}, {3, 2}, {3, 7}, {4, 5}, {4, 7}};
  for (int k = 0; k < n; ++k)
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        const int c[8] = {node3(i, j, k, n, n),         node3(i + 1, j, k, n, n),
                          node3(i + 1, j + 1, k, n, n), node3(i, j + 1, k, n, n),
                          node3(i, j, k + 1, n, n),     node3(i + 1, j, k + 1, n, n),
                          node3(i + 1, j + 1, k + 1, n, n), node3(i, j + 1, k + 1, n, n)};
        for (const auto &p : paths)
          for (int v : {c[0], c[p[0]], c[p[1]], c[6]}) m.conn.push_back(v);
      }
  return m;
}

int failures = 0, runs = 0;

void check(const Mesh &m, int epc, const std::vector<long long> *seg = nullptr, bool orient = true) {
  Mesh w = m;
  if (orient) {
    long long swapped = 0;
    if (femhot_orient_elements(w.et, w.nn(), w.xyz.data(), w.ne(), w.conn.data(), &swapped) != FEMHOT_OK) {
      std::printf("orient_elements failed: %s\n", femhot_last_error());
      ++failures;
      return;
    }
  }
  femhot_mesh_desc d{};
  d.elem_type = w.et;
  d.physics = FEMHOT_THERMAL;
  d.n_nodes = w.nn();
  d.n_elems = w.ne();
  d.coords = w.xyz.data();
  d.conn = w.conn.data();
  d.n_segments = seg ? (int)seg->size() - 1 : 1;
  d.segment_offsets = seg ? seg->data() : nullptr;
  long long st[8] = {0};
  const int rc = femhot_chunk_plan_check(&d, epc, st);
  ++runs;
  if (rc != FEMHOT_OK) {
    std::printf("FAIL et=%d ne=%lld epc=%d: %s\n", w.et, w.ne(), epc, femhot_last_error());
    ++failures;
  }
}

void shuffle_elements(Mesh &m, unsigned seed) {
  std::mt19937 rng(seed);
  const int n = m.npe();
  for (long long e = m.ne() - 1; e > 0; --e) {
    const long long f = (long long)(rng() % (unsigned)(e + 1));
    for (int a = 0; a < n; ++a) std::swap(m.conn[e * n + a], m.conn[f * n + a]);
  }
}

}  // namespace

int main() {
  const int epcs[] = {1, 2, 7, 32, 100, 256, 1024, 4096};
  std::vector<Mesh> ms = {quads(12), hexes(6, 5, 4), tets(5, 3u)};
  for (const Mesh &m : ms)
    for (int epc : epcs) check(m, epc);
  for (Mesh m : ms) {  // an element order with no locality: the planner cuts Morton chunks
    shuffle_elements(m, 7u);
    for (int epc : epcs) check(m, epc);
  }
  {  // uneven meshes, one segment each, the second one empty; a few nodes in no element
    Mesh all{FEMHOT_TET4, {}, {}};
    std::vector<long long> seg = {0};
    for (int n : {1, 4, 2, 6}) {
      Mesh t = tets(n, 10u + n);
      const int off = (int)all.nn();
      all.xyz.insert(all.xyz.end(), t.xyz.begin(), t.xyz.end());
      for (int v : t.conn) all.conn.push_back(v + off);
      seg.push_back(all.ne());
      if (seg.size() == 2) seg.push_back(all.ne());
    }
    for (int k = 0; k < 5; ++k)
      for (int d = 0; d < 3; ++d) all.xyz.push_back(50.0 + k + d);
    for (int epc : epcs) check(all, epc, &seg);
  }
  {  // a collapsed hexahedron (one node listed twice)
    Mesh h = hexes(3, 3, 3);
    h.conn[4 * 8 + 5] = h.conn[4 * 8 + 4];
    for (int epc : epcs) check(h, epc, nullptr, false);
  }
  std::printf("chunk plan checks: %d runs, %d failures\n", runs, failures);
  return failures ? 1 : 0;
}
