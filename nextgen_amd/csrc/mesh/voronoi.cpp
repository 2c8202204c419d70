// Periodic-box Voronoi meshes: the analytic uniform lattice and a clipping
// builder for jittered lattices.
//
// These stand in for Stitch's `PART ... SIMPLE_BOX ... PERIODIC_X/Y/Z` output
// (src/stitch/stitch.cpp, written by StitchNS.cpp:8122 writeMles): cells are
// Voronoi regions of the generator points, x_cv is the volume centroid
// (StitchNS.cpp:5008-5015 returns x_cv "back to a true centroid"), x_fa the face
// centroid and n_fa the face area vector.
//
// Jittered cells are built by half-space clipping: start from a box around the
// generator and intersect it with the bisector half-space of every candidate
// neighbour, nearest first, stopping once a candidate is farther than twice the
// largest vertex radius (no farther bisector can cut the cell).
#include "mesh.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <thread>

namespace cgmesh {

uint64_t morton3(uint32_t i, uint32_t j, uint32_t k) {
  auto spread = [](uint64_t x) {
    x &= 0x1fffff;
    x = (x | x << 32) & 0x1f00000000ffffULL;
    x = (x | x << 16) & 0x1f0000ff0000ffULL;
    x = (x | x << 8) & 0x100f00f00f00f00fULL;
    x = (x | x << 4) & 0x10c30c30c30c30c3ULL;
    x = (x | x << 2) & 0x1249249249249249ULL;
    return x;
  };
  return spread(i) | (spread(j) << 1) | (spread(k) << 2);
}

double hash_uniform(uint64_t seed, uint64_t counter) {
  uint64_t z = seed * 0x9E3779B97F4A7C15ULL + counter * 0xD1B54A32D192ED03ULL + 0x632BE59BD9B4E019ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  z = z ^ (z >> 31);
  return double(z >> 11) * (1.0 / 9007199254740992.0);
}

int default_threads() {
  unsigned n = std::thread::hardware_concurrency();
  if (n == 0) n = 1;
  if (n > 16) n = 16;
  return int(n);
}

namespace {

// Lattice cells in Morton order: owner index of lattice cell (i,j,k).
struct MortonLattice {
  int n[3];
  std::vector<int> owner;      // [(i*ny+j)*nz+k] -> owner index
  std::vector<int> ijk;        // [owner][3]
  void build(int nx, int ny, int nz) {
    n[0] = nx; n[1] = ny; n[2] = nz;
    const int ncv = nx * ny * nz;
    std::vector<std::pair<uint64_t, int>> key(ncv);
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j)
        for (int k = 0; k < nz; ++k) {
          const int lin = (i * ny + j) * nz + k;
          key[lin] = {morton3(i, j, k), lin};
        }
    std::sort(key.begin(), key.end());
    owner.assign(ncv, -1);
    ijk.assign(3 * ncv, 0);
    for (int r = 0; r < ncv; ++r) {
      const int lin = key[r].second;
      owner[lin] = r;
      ijk[3 * r + 0] = lin / (ny * nz);
      ijk[3 * r + 1] = (lin / nz) % ny;
      ijk[3 * r + 2] = lin % nz;
    }
  }
  // owner and image shift of lattice index (possibly outside [0,n))
  inline int wrap(int i, int j, int k, int8_t s[3]) const {
    const int a[3] = {i, j, k};
    int w[3];
    for (int d = 0; d < 3; ++d) {
      int q = a[d] >= 0 ? a[d] / n[d] : -((-a[d] + n[d] - 1) / n[d]);
      w[d] = a[d] - q * n[d];
      s[d] = int8_t(q);
    }
    return owner[(w[0] * n[1] + w[1]) * n[2] + w[2]];
  }
};

struct V3 {
  double x, y, z;
};
inline V3 operator-(const V3& a, const V3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator+(const V3& a, const V3& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator*(double s, const V3& a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dot(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(const V3& a, const V3& b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// Convex polyhedron as vertex loops (counter-clockwise seen from outside).
struct Poly {
  std::vector<V3> v;
  std::vector<std::vector<int>> loop;
  std::vector<int> id;  // clipping-plane id (>= 0 candidate, < 0 initial box)

  void box(double bx, double by, double bz) {
    v = {{-bx, -by, -bz}, {bx, -by, -bz}, {bx, by, -bz}, {-bx, by, -bz},
         {-bx, -by, bz},  {bx, -by, bz},  {bx, by, bz},  {-bx, by, bz}};
    loop = {{0, 3, 2, 1}, {4, 5, 6, 7}, {0, 1, 5, 4}, {2, 3, 7, 6}, {0, 4, 7, 3}, {1, 2, 6, 5}};
    id = {-1, -2, -3, -4, -5, -6};
  }

  double rmax2() const {
    double r = 0.0;
    for (const auto& l : loop)
      for (int a : l) r = std::max(r, dot(v[a], v[a]));
    return r;
  }

  // Keep {x : x.r <= rr2}. Returns 1 if the cell changed, 0 if not, -1 on a
  // topological failure.
  int clip(const V3& r, double rr2, double eps, int pid) {
    const int nv = int(v.size());
    std::vector<double> s(nv);
    bool any_out = false, any_in = false;
    for (int a = 0; a < nv; ++a) {
      s[a] = dot(v[a], r) - rr2;
    }
    for (const auto& l : loop)
      for (int a : l) {
        if (s[a] > eps) any_out = true; else any_in = true;
      }
    if (!any_out) return 0;
    if (!any_in) return -1;
    struct Edge { int a, b, x; };
    std::vector<Edge> edges;
    auto edge_vertex = [&](int a, int b) {
      const int lo = std::min(a, b), hi = std::max(a, b);
      for (const auto& e : edges)
        if (e.a == lo && e.b == hi) return e.x;
      // interpolate from the inside vertex toward the outside one
      const int in = s[lo] > eps ? hi : lo;
      const int out = in == lo ? hi : lo;
      const double t = -s[in] / (s[out] - s[in]);
      v.push_back(v[in] + t * (v[out] - v[in]));
      const int x = int(v.size()) - 1;
      edges.push_back({lo, hi, x});
      return x;
    };
    std::vector<std::vector<int>> nloop;
    std::vector<int> nid;
    std::vector<std::pair<int, int>> cap;  // entry -> exit along the cap plane
    for (size_t f = 0; f < loop.size(); ++f) {
      const auto& l = loop[f];
      const int n = int(l.size());
      std::vector<int> out;
      int exit_pt = -1, entry_pt = -1, nexit = 0;
      for (int k = 0; k < n; ++k) {
        const int a = l[k], b = l[(k + 1) % n];
        const bool ina = s[a] <= eps, inb = s[b] <= eps;
        if (ina) out.push_back(a);
        if (ina != inb) {
          const int x = edge_vertex(a, b);
          out.push_back(x);
          if (ina) { exit_pt = x; ++nexit; } else { entry_pt = x; }
        }
      }
      if (nexit > 1) return -1;
      if (exit_pt >= 0 && entry_pt >= 0) cap.push_back({entry_pt, exit_pt});
      if (out.size() >= 3) {
        nloop.push_back(out);
        nid.push_back(id[f]);
      }
    }
    if (cap.size() >= 3) {
      // chain entry->exit; an exit of one face is the entry of its neighbour
      std::vector<int> c;
      c.reserve(cap.size());
      int cur = cap[0].first;
      for (size_t it = 0; it < cap.size(); ++it) {
        int nxt = -1;
        for (const auto& e : cap)
          if (e.first == cur) { nxt = e.second; break; }
        if (nxt < 0) return -1;
        c.push_back(cur);
        cur = nxt;
      }
      if (cur != cap[0].first) return -1;
      nloop.push_back(c);
      nid.push_back(pid);
    }
    loop.swap(nloop);
    id.swap(nid);
    // compact the vertex array
    std::vector<int> remap(v.size(), -1);
    std::vector<V3> nvv;
    for (auto& l : loop)
      for (int& a : l) {
        if (remap[a] < 0) { remap[a] = int(nvv.size()); nvv.push_back(v[a]); }
        a = remap[a];
      }
    v.swap(nvv);
    return 1;
  }
};

struct HalfFace {
  int nbr;
  int8_t s[3];
  double n[3];  // outward area vector
  double x[3];  // centroid, generator-relative
};

struct BoundaryFace {
  int zone;
  double n[3];  // outward area vector
  double x[3];  // centroid, generator-relative
};

struct CellOut {
  double vol, xc[3], r2;
  std::vector<HalfFace> hf;
  std::vector<BoundaryFace> bf;
};

// clipping-plane ids of the channel walls y = 0 and y = Ly (candidate ids are
// their index in the candidate list, the initial box's faces are negative)
constexpr int kWallId = 1 << 28;

std::string sort_and_check(Voronoi& m) {
  // orient faces cv0 < cv1 and sort by (cv0, cv1, shift)
  const int nfa = m.nfa;
  std::vector<int> order(nfa);
  for (int f = 0; f < nfa; ++f) {
    if (m.cvofa[2 * f] == m.cvofa[2 * f + 1])
      return "face joins a cell to its own periodic image (domain too thin)";
    if (m.cvofa[2 * f] > m.cvofa[2 * f + 1]) {
      const int a = m.cvofa[2 * f], b = m.cvofa[2 * f + 1];
      for (int d = 0; d < 3; ++d) {
        // move x_fa into the new cv0's frame
        m.x_fa[3 * f + d] -= m.shift[3 * f + d] * m.L[d];
        m.shift[3 * f + d] = int8_t(-m.shift[3 * f + d]);
        m.n_fa[3 * f + d] = -m.n_fa[3 * f + d];
      }
      m.cvofa[2 * f] = b;
      m.cvofa[2 * f + 1] = a;
    }
    order[f] = f;
  }
  auto key = [&](int f) {
    return std::make_tuple(m.cvofa[2 * f], m.cvofa[2 * f + 1], m.shift[3 * f], m.shift[3 * f + 1],
                           m.shift[3 * f + 2]);
  };
  std::sort(order.begin(), order.end(), [&](int a, int b) { return key(a) < key(b); });
  for (int i = 1; i < nfa; ++i)
    if (key(order[i]) == key(order[i - 1])) return "duplicate face";
  Voronoi t = m;
  for (int i = 0; i < nfa; ++i) {
    const int f = order[i];
    for (int d = 0; d < 2; ++d) m.cvofa[2 * i + d] = t.cvofa[2 * f + d];
    for (int d = 0; d < 3; ++d) {
      m.shift[3 * i + d] = t.shift[3 * f + d];
      m.n_fa[3 * i + d] = t.n_fa[3 * f + d];
      m.x_fa[3 * i + d] = t.x_fa[3 * f + d];
    }
  }
  return "";
}

}  // namespace

std::string build_lattice(Voronoi& m, int nx, int ny, int nz, double Lx, double Ly, double Lz) {
  if (nx < 5 || ny < 5 || nz < 5) return "lattice needs at least 5 cells per direction";
  MortonLattice ml;
  ml.build(nx, ny, nz);
  m = Voronoi();
  m.ncv = nx * ny * nz;
  m.L[0] = Lx; m.L[1] = Ly; m.L[2] = Lz;
  m.lattice[0] = nx; m.lattice[1] = ny; m.lattice[2] = nz;
  const double h[3] = {Lx / nx, Ly / ny, Lz / nz};
  const double vol = h[0] * h[1] * h[2];
  const double area[3] = {h[1] * h[2], h[0] * h[2], h[0] * h[1]};
  m.x_vv.resize(3 * m.ncv);
  m.vol.assign(m.ncv, vol);
  for (int c = 0; c < m.ncv; ++c)
    for (int d = 0; d < 3; ++d) m.x_vv[3 * c + d] = (ml.ijk[3 * c + d] + 0.5) * h[d];
  m.x_cv = m.x_vv;
  // every vertex of a box cell is half its diagonal from the generator
  m.r_vv.assign(m.ncv, std::sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]) / 2.0);
  m.nfa = 3 * m.ncv;
  m.cvofa.resize(2 * m.nfa);
  m.shift.assign(3 * m.nfa, 0);
  m.n_fa.assign(3 * m.nfa, 0.0);
  m.x_fa.resize(3 * m.nfa);
  int f = 0;
  for (int c = 0; c < m.ncv; ++c) {
    for (int d = 0; d < 3; ++d) {
      int q[3] = {ml.ijk[3 * c], ml.ijk[3 * c + 1], ml.ijk[3 * c + 2]};
      q[d] += 1;
      int8_t s[3];
      const int nb = ml.wrap(q[0], q[1], q[2], s);
      m.cvofa[2 * f] = c;
      m.cvofa[2 * f + 1] = nb;
      for (int e = 0; e < 3; ++e) {
        m.shift[3 * f + e] = s[e];
        m.x_fa[3 * f + e] = m.x_cv[3 * c + e];
      }
      m.x_fa[3 * f + d] = m.x_cv[3 * c + d] + 0.5 * h[d];
      m.n_fa[3 * f + d] = area[d];
      ++f;
    }
  }
  return sort_and_check(m);
}

namespace {
std::string build_clipped(Voronoi& m, int nx, int ny, int nz, double Lx, double Ly, double Lz, double jitter,
                          uint64_t seed, int nthreads, bool walls_y) {
  if (nx < 8 || ny < 8 || nz < 8) return "jittered mesh needs at least 8 cells per direction";
  if (!(jitter >= 0.0 && jitter < 0.5)) return "jitter must be in [0, 0.5)";
  MortonLattice ml;
  ml.build(nx, ny, nz);
  m = Voronoi();
  const int ncv = nx * ny * nz;
  m.ncv = ncv;
  m.L[0] = Lx; m.L[1] = Ly; m.L[2] = Lz;
  m.lattice[0] = nx; m.lattice[1] = ny; m.lattice[2] = nz;
  const double h[3] = {Lx / nx, Ly / ny, Lz / nz};
  const double hmax = std::max(h[0], std::max(h[1], h[2]));
  m.x_vv.resize(3 * ncv);
  for (int c = 0; c < ncv; ++c) {
    const int* q = &ml.ijk[3 * c];
    const uint64_t lin = (uint64_t(q[0]) * ny + q[1]) * nz + q[2];
    for (int d = 0; d < 3; ++d) {
      const double u = hash_uniform(seed, 3 * lin + d);
      m.x_vv[3 * c + d] = (q[d] + 0.5 + jitter * (2.0 * u - 1.0)) * h[d];
    }
  }

  std::vector<CellOut> cells(ncv);
  std::vector<std::string> errs(nthreads);
  auto work = [&](int t) {
    Poly P;
    struct Cand { double d2; V3 r; int nbr; int8_t s[3]; };
    std::vector<Cand> cand;
    for (int c = t; c < ncv; c += nthreads) {
      const V3 p = {m.x_vv[3 * c], m.x_vv[3 * c + 1], m.x_vv[3 * c + 2]};
      const int* q = &ml.ijk[3 * c];
      bool done = false;
      for (int R = 2; R <= 4 && !done; ++R) {
        cand.clear();
        for (int di = -R; di <= R; ++di)
          for (int dj = -R; dj <= R; ++dj)
            for (int dk = -R; dk <= R; ++dk) {
              if (!di && !dj && !dk) continue;
              if (walls_y && (q[1] + dj < 0 || q[1] + dj >= ny)) continue;  // no images across the walls
              Cand cc;
              cc.nbr = ml.wrap(q[0] + di, q[1] + dj, q[2] + dk, cc.s);
              const V3 x = {m.x_vv[3 * cc.nbr] + cc.s[0] * Lx, m.x_vv[3 * cc.nbr + 1] + cc.s[1] * Ly,
                            m.x_vv[3 * cc.nbr + 2] + cc.s[2] * Lz};
              cc.r = x - p;
              cc.d2 = dot(cc.r, cc.r);
              cand.push_back(cc);
            }
        std::sort(cand.begin(), cand.end(), [](const Cand& a, const Cand& b) { return a.d2 < b.d2; });
        P.box(3.0 * hmax, 3.0 * hmax, 3.0 * hmax);
        const double eps = 1e-13 * hmax * hmax;
        bool exhausted = true, bad = false;
        if (walls_y) {
          // keep y >= 0 and y <= Ly (generator-relative half-spaces)
          if (P.clip({0.0, -1.0, 0.0}, p.y, 1e-13 * hmax, kWallId) < 0 ||
              P.clip({0.0, 1.0, 0.0}, Ly - p.y, 1e-13 * hmax, kWallId + 1) < 0)
            bad = true;
        }
        double r2 = P.rmax2();
        for (size_t ic = 0; ic < cand.size(); ++ic) {
          if (0.25 * cand[ic].d2 > r2) { exhausted = false; break; }
          const int rc = P.clip(cand[ic].r, 0.5 * cand[ic].d2, eps, int(ic));
          if (rc < 0) { bad = true; break; }
          if (rc > 0) r2 = P.rmax2();
        }
        if (bad) { errs[t] = "clipping failed at cell " + std::to_string(c); return; }
        if (exhausted) continue;  // widen the candidate block
        for (int id : P.id)
          if (id < 0) { errs[t] = "cell not closed by neighbours at cell " + std::to_string(c); return; }
        // geometry relative to the generator
        CellOut& co = cells[c];
        double vol = 0.0, xc[3] = {0, 0, 0};
        co.hf.clear();
        for (size_t f = 0; f < P.loop.size(); ++f) {
          const auto& l = P.loop[f];
          V3 A = {0, 0, 0};
          for (size_t k = 1; k + 1 < l.size(); ++k)
            A = A + 0.5 * cross(P.v[l[k]] - P.v[l[0]], P.v[l[k + 1]] - P.v[l[0]]);
          const double Am = std::sqrt(dot(A, A));
          V3 xf = {0, 0, 0};
          double wsum = 0.0;
          for (size_t k = 1; k + 1 < l.size(); ++k) {
            const V3 a = P.v[l[0]], b = P.v[l[k]], e = P.v[l[k + 1]];
            const double w = Am > 0 ? dot(cross(b - a, e - a), A) / Am : 0.0;
            xf = xf + (w / 3.0) * (a + b + e);
            wsum += w;
            const double tv = dot(a, cross(b, e)) / 6.0;
            vol += tv;
            const V3 tc = 0.25 * (a + b + e);
            xc[0] += tv * tc.x; xc[1] += tv * tc.y; xc[2] += tv * tc.z;
          }
          if (wsum > 0) xf = (1.0 / wsum) * xf;
          else {
            xf = {0, 0, 0};
            for (int a : l) xf = xf + (1.0 / l.size()) * P.v[a];
          }
          if (P.id[f] >= kWallId) {
            BoundaryFace b;
            b.zone = P.id[f] - kWallId;
            b.n[0] = A.x; b.n[1] = A.y; b.n[2] = A.z;
            b.x[0] = xf.x; b.x[1] = xf.y; b.x[2] = xf.z;
            co.bf.push_back(b);
            continue;
          }
          HalfFace hf;
          const Cand& cc = cand[P.id[f]];
          hf.nbr = cc.nbr;
          for (int d = 0; d < 3; ++d) hf.s[d] = cc.s[d];
          hf.n[0] = A.x; hf.n[1] = A.y; hf.n[2] = A.z;
          hf.x[0] = xf.x; hf.x[1] = xf.y; hf.x[2] = xf.z;
          co.hf.push_back(hf);
        }
        co.vol = vol;
        co.r2 = P.rmax2();
        for (int d = 0; d < 3; ++d) co.xc[d] = xc[d] / vol;
        std::sort(co.hf.begin(), co.hf.end(), [](const HalfFace& a, const HalfFace& b) {
          return std::make_tuple(a.nbr, a.s[0], a.s[1], a.s[2]) < std::make_tuple(b.nbr, b.s[0], b.s[1], b.s[2]);
        });
        done = true;
      }
      if (!done) { errs[t] = "candidate block too small at cell " + std::to_string(c); return; }
    }
  };
  {
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; ++t) th.emplace_back(work, t);
    for (auto& x : th) x.join();
  }
  for (auto& e : errs)
    if (!e.empty()) return e;

  m.x_cv.resize(3 * ncv);
  m.vol.resize(ncv);
  m.r_vv.resize(ncv);
  for (int c = 0; c < ncv; ++c) {
    m.vol[c] = cells[c].vol;
    m.r_vv[c] = std::sqrt(cells[c].r2);
    for (int d = 0; d < 3; ++d) m.x_cv[3 * c + d] = m.x_vv[3 * c + d] + cells[c].xc[d];
  }
  // merge the two sides of every face: the lower cell index owns the geometry
  const double tiny = 1e-13 * hmax * hmax, sliver = 1e-9 * hmax * hmax;
  auto find = [&](int c, int nbr, const int8_t s[3]) -> const HalfFace* {
    for (const auto& hf : cells[c].hf)
      if (hf.nbr == nbr && hf.s[0] == s[0] && hf.s[1] == s[1] && hf.s[2] == s[2]) return &hf;
    return nullptr;
  };
  int ndrop = 0;
  for (int c = 0; c < ncv; ++c) {
    for (const auto& hf : cells[c].hf) {
      if (hf.nbr == c) return "cell touches its own periodic image";
      const int8_t ms[3] = {int8_t(-hf.s[0]), int8_t(-hf.s[1]), int8_t(-hf.s[2])};
      const HalfFace* other = find(hf.nbr, c, ms);
      const double a = std::sqrt(hf.n[0] * hf.n[0] + hf.n[1] * hf.n[1] + hf.n[2] * hf.n[2]);
      if (!other) {
        if (a > sliver) return "unmatched face of area " + std::to_string(a) + " at cell " + std::to_string(c);
        ++ndrop;
        continue;
      }
      if (hf.nbr < c) continue;
      if (a < tiny) { ++ndrop; continue; }
      m.cvofa.push_back(c);
      m.cvofa.push_back(hf.nbr);
      for (int d = 0; d < 3; ++d) {
        m.shift.push_back(hf.s[d]);
        m.n_fa.push_back(hf.n[d]);
        m.x_fa.push_back(m.x_vv[3 * c + d] + hf.x[d]);
      }
    }
  }
  m.nfa = int(m.cvofa.size() / 2);
  m.ndropped = ndrop;
  if (walls_y) {
    // boundary faces zone by zone, each zone's faces by cell
    m.bf_zone_name = {"y0", "y1"};
    m.bf_zone_i.assign(3, 0);
    for (int z = 0; z < 2; ++z) {
      for (int c = 0; c < ncv; ++c)
        for (const auto& b : cells[c].bf) {
          if (b.zone != z) continue;
          m.cvobf.push_back(c);
          for (int d = 0; d < 3; ++d) {
            m.n_bf.push_back(b.n[d]);
            m.x_bf.push_back(m.x_vv[3 * c + d] + b.x[d]);
          }
        }
      m.bf_zone_i[z + 1] = int(m.cvobf.size());
    }
    m.nbf = int(m.cvobf.size());
  }
  return sort_and_check(m);
}
}  // namespace

std::string build_jittered(Voronoi& m, int nx, int ny, int nz, double Lx, double Ly, double Lz,
                           double jitter, uint64_t seed, int nthreads) {
  return build_clipped(m, nx, ny, nz, Lx, Ly, Lz, jitter, seed, nthreads, false);
}

// the domain extents of a box [0, L]: its corner node (Lx, Ly, Lz) holds every maximum
static void set_box_x_max(Voronoi& m) {
  const double* L = m.L;
  m.x_max[0] = L[0];
  m.x_max[1] = L[1];
  m.x_max[2] = L[2];
  m.x_max[3] = std::sqrt(L[1] * L[1] + L[2] * L[2]);
  m.x_max[4] = std::sqrt(L[0] * L[0] + L[1] * L[1]);
  m.has_x_max = true;
  m.x_ext_nodes.assign(L, L + 3);
}

std::string build_channel(Voronoi& m, int nx, int ny, int nz, double Lx, double Ly, double Lz, double jitter,
                          uint64_t seed, int nthreads) {
  if (jitter > 0.0) {
    std::string e = build_clipped(m, nx, ny, nz, Lx, Ly, Lz, jitter, seed, nthreads, true);
    if (e.empty()) set_box_x_max(m);
    return e;
  }
  // the uniform lattice of box cells: periodic faces in x and z, interior faces
  // in y, a wall face on each cell of the first and last layer
  std::string e = build_lattice(m, nx, ny, nz, Lx, Ly, Lz);
  if (!e.empty()) return e;
  const double h[3] = {Lx / nx, Ly / ny, Lz / nz};
  Voronoi t = m;
  m.cvofa.clear(); m.shift.clear(); m.n_fa.clear(); m.x_fa.clear();
  for (int f = 0; f < t.nfa; ++f) {
    if (t.shift[3 * f + 1] != 0) continue;  // the wrap across y becomes two walls
    for (int d = 0; d < 2; ++d) m.cvofa.push_back(t.cvofa[2 * f + d]);
    for (int d = 0; d < 3; ++d) {
      m.shift.push_back(t.shift[3 * f + d]);
      m.n_fa.push_back(t.n_fa[3 * f + d]);
      m.x_fa.push_back(t.x_fa[3 * f + d]);
    }
  }
  m.nfa = int(m.cvofa.size() / 2);
  m.bf_zone_name = {"y0", "y1"};
  m.bf_zone_i.assign(3, 0);
  for (int z = 0; z < 2; ++z) {
    for (int c = 0; c < m.ncv; ++c) {
      const double y = m.x_cv[3 * c + 1];
      const bool first = y < h[1], last = y > Ly - h[1];
      if ((z == 0 && !first) || (z == 1 && !last)) continue;
      m.cvobf.push_back(c);
      const double sgn = z == 0 ? -1.0 : 1.0;
      m.n_bf.push_back(0.0); m.n_bf.push_back(sgn * h[0] * h[2]); m.n_bf.push_back(0.0);
      m.x_bf.push_back(m.x_cv[3 * c]);
      m.x_bf.push_back(z == 0 ? 0.0 : Ly);
      m.x_bf.push_back(m.x_cv[3 * c + 2]);
    }
    m.bf_zone_i[z + 1] = int(m.cvobf.size());
  }
  m.nbf = int(m.cvobf.size());
  set_box_x_max(m);
  return "";
}

std::string build_from_arrays(Voronoi& m, int ncv, const double* x_cv, const double* vol, int nfa,
                              const int* cvofa, const int* shift, const double* n_fa,
                              const double* x_fa, const double L[3]) {
  if (ncv <= 0 || nfa <= 0) return "empty mesh";
  m = Voronoi();
  m.ncv = ncv;
  for (int d = 0; d < 3; ++d) m.L[d] = L[d];
  m.x_cv.assign(x_cv, x_cv + 3 * ncv);
  m.x_vv = m.x_cv;
  m.vol.assign(vol, vol + ncv);
  m.nfa = nfa;
  m.cvofa.assign(cvofa, cvofa + 2 * nfa);
  m.shift.resize(3 * nfa);
  for (int i = 0; i < 3 * nfa; ++i) {
    if (shift[i] < -1 || shift[i] > 1) return "periodic shift must be -1, 0 or 1";
    m.shift[i] = int8_t(shift[i]);
  }
  m.n_fa.assign(n_fa, n_fa + 3 * nfa);
  m.x_fa.assign(x_fa, x_fa + 3 * nfa);
  for (int f = 0; f < nfa; ++f)
    for (int s = 0; s < 2; ++s)
      if (cvofa[2 * f + s] < 0 || cvofa[2 * f + s] >= ncv) return "cvofa index out of range";
  for (int c = 0; c < ncv; ++c)
    if (!(vol[c] > 0.0)) return "non-positive cell volume";
  return sort_and_check(m);
}

}  // namespace cgmesh
