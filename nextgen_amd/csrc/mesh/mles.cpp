// .mles mesh files and .sles data files: the record format the reference reads in
// StripedMesh::read_mles_ (src/core/common/StripedMesh.hpp:1336-2400) and writes in
// StitchNS::writeMles (src/stitch/StitchNS.cpp:8122) and StaticSolver::writeData
// (src/core/common/StaticSolver.cpp:7560).
//
// A file is two int32 (magic number, io version) followed by records. Each record
// is a 256-byte header -- name[52], int32 id, int64 skip (bytes from this header to
// the next), int32 idata[16], double rdata[12], uint64 ui8data[4]
// (src/core/common/Defs.hpp:200,228-249, MpiStuff.cpp:1225-1266) -- and its data.
// 64-bit counts in idata are split as (value - (msw << 31), msw = value >> 31)
// (ByteSwap.hpp:117-140). Files written on a machine of the other byte order are
// recognised by a byte-swapped magic number and swapped on reading.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "mesh.hpp"

namespace cgmesh {

namespace {

struct RecHeader {
  char name[52];
  int32_t id;
  int64_t skip;
  int32_t idata[16];
  double rdata[12];
  uint64_t ui8[4];
};
static_assert(sizeof(RecHeader) == 256, "record header is 256 bytes");

// record ids (Defs.hpp:27-185)
constexpr int32_t kMagic = 123581321;
constexpr int32_t kCounts = 10001;      // UGP_IO_NO_FA_BF_CV_COUNTS
constexpr int32_t kI0 = 11;             // UGP_IO_I0
constexpr int32_t kD0 = 12;             // UGP_IO_D0
constexpr int32_t kFaD2 = 25;           // UGP_IO_FA_D2
constexpr int32_t kXNo = 31;            // UGP_IO_X_NO
constexpr int32_t kPbiPno = 36;         // UGP_IO_PBI_PNO
constexpr int32_t kCvD1 = 42;           // UGP_IO_CV_D1
constexpr int32_t kCvD2 = 43;           // UGP_IO_CV_D2
constexpr int32_t kCvZoneHeader = 46;   // UGP_IO_CV_ZONE_HEADER
constexpr int32_t kEof = 51;            // UGP_IO_EOF
constexpr int32_t kBfZoneHeader = 777;  // UGP_IO_BF_ZONE_HEADER
constexpr int32_t kCvobf = 2004;        // UGP_IO_CVOBF_INT8
constexpr int32_t kNoobf = 2005;        // UGP_IO_NOOBF_I_AND_V_INT8
constexpr int32_t kCvofa = 2006;        // UGP_IO_CVOFA_INT8
constexpr int32_t kNoofa = 2007;        // UGP_IO_NOOFA_I_AND_V_INT8
constexpr int32_t kBfD1 = 2120;         // UGP_IO_BF_D1
constexpr int32_t kBfD2 = 2121;         // UGP_IO_BF_D2
constexpr int32_t kBfDN33 = 2122;       // UGP_IO_BF_DN33
constexpr int32_t kPeriodic = 4123;     // UGP_IO_PERIODIC_TRANSFORM
constexpr int32_t kZoneFluid = 1;       // CV_ZONE_FLUID
constexpr int32_t kTransformCart = 1;   // PERIODIC_TRANSFORM_CART

template <class T>
void swap_bytes(T* p, size_t n) {
  for (size_t i = 0; i < n; ++i) {
    unsigned char* b = reinterpret_cast<unsigned char*>(p + i);
    std::reverse(b, b + sizeof(T));
  }
}

void swap_header(RecHeader& h) {
  swap_bytes(&h.id, 1);
  swap_bytes(&h.skip, 1);
  swap_bytes(h.idata, 16);
  swap_bytes(h.rdata, 12);
  swap_bytes(h.ui8, 4);
}

// 64-bit count from an idata pair; a negative low word with a zero high word is
// the legacy form of a count between 2^31 and 2^32 (ByteSwap.hpp:117-131)
bool get_count(const int32_t* p, int64_t& v) {
  int64_t lsw = p[0], msw = p[1];
  if (lsw < 0) {
    if (msw != 0) return false;
    lsw += 4294967296LL;
  }
  v = (msw << 31) + lsw;
  return v >= 0;
}

void set_count(int32_t* p, int64_t v) {
  const int64_t msw = v >> 31;
  p[0] = int32_t(v - (msw << 31));
  p[1] = int32_t(msw);
}

// periodic bits of face zones 0..25 (zone 26 holds the interior faces): bit 2i
// applies transform i, bit 2i+1 its inverse (StripedMesh.hpp:3043-3105)
constexpr int kZoneBits[26] = {
    1 << 0,                       1 << 1,                       1 << 2,
    1 << 3,                       (1 << 2) | (1 << 0),          (1 << 2) | (1 << 1),
    (1 << 3) | (1 << 0),          (1 << 3) | (1 << 1),          1 << 4,
    1 << 5,                       (1 << 4) | (1 << 0),          (1 << 4) | (1 << 1),
    (1 << 5) | (1 << 0),          (1 << 5) | (1 << 1),          (1 << 4) | (1 << 2),
    (1 << 4) | (1 << 3),          (1 << 5) | (1 << 2),          (1 << 5) | (1 << 3),
    (1 << 4) | (1 << 2) | (1 << 0), (1 << 4) | (1 << 2) | (1 << 1), (1 << 4) | (1 << 3) | (1 << 0),
    (1 << 4) | (1 << 3) | (1 << 1), (1 << 5) | (1 << 2) | (1 << 0), (1 << 5) | (1 << 2) | (1 << 1),
    (1 << 5) | (1 << 3) | (1 << 0), (1 << 5) | (1 << 3) | (1 << 1)};

class Reader {
 public:
  explicit Reader(const std::string& path) { fp_ = std::fopen(path.c_str(), "rb"); }
  ~Reader() {
    if (fp_) std::fclose(fp_);
  }
  bool ok() const { return fp_ != nullptr; }
  int64_t size() {
    std::fseek(fp_, 0, SEEK_END);
    return int64_t(std::ftell(fp_));
  }
  template <class T>
  bool read(int64_t off, T* dst, size_t n, bool swap) {
    if (n == 0) return true;
    if (std::fseek(fp_, long(off), SEEK_SET) != 0) return false;
    if (std::fread(dst, sizeof(T), n, fp_) != n) return false;
    if (swap) swap_bytes(dst, n);
    return true;
  }

 private:
  FILE* fp_ = nullptr;
};

std::string hname(const RecHeader& h) {
  char buf[53];
  std::memcpy(buf, h.name, 52);
  buf[52] = 0;
  return std::string(buf);
}

}  // namespace

int64_t Mles::nfa_p() const {
  int64_t s = 0;
  for (int z = 0; z < 26; ++z) s += nfa_zone[z];
  return s;
}

std::string read_mles(const std::string& path, Mles& f) {
  f = Mles();
  Reader r(path);
  if (!r.ok()) return "cannot open " + path;
  const int64_t fsize = r.size();
  int32_t head[2] = {0, 0};
  if (!r.read(0, head, 2, false)) return path + ": too short for a record file";
  bool swap = false;
  if (head[0] != kMagic) {
    swap_bytes(head, 2);
    if (head[0] != kMagic) return path + ": not a mesh record file (magic number)";
    swap = true;
  }
  f.swapped = swap;
  f.version = head[1];
  if (f.version == 3)
    return path + ": io version 3 mesh files are not supported (io version 5 is what Stitch writes)";
  if (f.version != 5) return path + ": unknown io version " + std::to_string(f.version);

  bool have_counts = false, done = false;
  int64_t off = 8, nbf_check = 0, ncv_check = 0;
  auto bad = [&](const std::string& what) { return path + ": " + what; };
  while (!done) {
    RecHeader h;
    if (off + 256 > fsize || !r.read(off, &h, 1, false)) return bad("truncated (no EOF record)");
    if (swap) swap_header(h);
    if (h.skip < 256 || off + h.skip > fsize) {
      if (h.id != kEof) return bad("record \"" + hname(h) + "\" runs past the end of the file");
    }
    const std::string name = hname(h);
    const int64_t data = off + 256;
    const int64_t avail = h.skip - 256;
    int64_t n = 0;
    auto need = [&](int64_t bytes) { return bytes >= 0 && bytes <= avail; };
    auto counted = [&](int64_t expect) {
      return get_count(h.idata, n) && n == expect;
    };
    switch (h.id) {
      case kCounts: {
        if (!get_count(h.idata + 0, f.nno) || !get_count(h.idata + 2, f.nbf) || !get_count(h.idata + 4, f.nfa) ||
            !get_count(h.idata + 6, f.ncv) || !get_count(h.idata + 8, f.nno_p) || !get_count(h.idata + 10, f.nno_pb))
          return bad("bad counts record");
        f.surface = h.idata[12] == 1;
        if (!need(2 * 27 * 8) || !r.read(data, f.nfa_zone, 27, swap) || !r.read(data + 27 * 8, f.noofa_zone, 27, swap))
          return bad("bad zone counts");
        int64_t s = 0;
        for (int z = 0; z < 27; ++z) {
          if (f.nfa_zone[z] < 0) return bad("negative zone count");
          s += f.nfa_zone[z];
        }
        if (s != f.nfa) return bad("face zone counts do not add up to nfa");
        have_counts = true;
        break;
      }
      case kPeriodic: {
        f.npt = h.idata[0];
        if (f.npt < 1 || f.npt > 3) return bad("periodic transform count must be 1..3");
        for (int i = 0; i < f.npt; ++i) {
          f.pt_kind[i] = h.idata[1 + i];
          for (int d = 0; d < 3; ++d) f.pt_data[i][d] = h.rdata[3 * i + d];
        }
        break;
      }
      case kBfZoneHeader: {
        MlesZone z;
        z.name = name;
        if (!get_count(h.idata + 2, z.nbf) || !get_count(h.idata + 4, z.noobf) || !get_count(h.idata + 6, z.ibf_f))
          return bad("bad boundary zone header");
        z.area = h.rdata[0];
        z.area_over_delta = h.rdata[1];
        for (int d = 0; d < 3; ++d) {
          z.n[d] = h.rdata[2 + d];
          z.x[d] = h.rdata[5 + d];
        }
        nbf_check += z.nbf;
        f.bf_zones.push_back(z);
        break;
      }
      case kCvZoneHeader: {
        int64_t nz = 0;
        if (!get_count(h.idata + 2, nz)) return bad("bad cell zone header");
        f.cv_zones.push_back(name);
        ncv_check += nz;
        break;
      }
      case kCvobf:
        if (!have_counts || !counted(f.nbf) || !need(8 * f.nbf)) return bad("bad cvobf record");
        f.cvobf.resize(size_t(f.nbf));
        if (!r.read(data, f.cvobf.data(), f.cvobf.size(), swap)) return bad("cannot read cvobf");
        break;
      case kCvofa:
        if (!have_counts || !counted(f.nfa) || h.idata[2] != 2 || !need(16 * f.nfa)) return bad("bad cvofa record");
        f.cvofa.resize(size_t(2 * f.nfa));
        if (!r.read(data, f.cvofa.data(), f.cvofa.size(), swap)) return bad("cannot read cvofa");
        break;
      case kNoofa:
      case kNoobf: {
        const bool fa = h.id == kNoofa;
        const int64_t nf = fa ? f.nfa : f.nbf;
        int64_t nv = 0;
        if (!have_counts || !counted(nf) || !get_count(h.idata + 2, nv) || !need(4 * nf + 8 * nv))
          return bad(fa ? "bad noofa record" : "bad noobf record");
        std::vector<int32_t> cnt(static_cast<size_t>(nf));
        std::vector<int64_t>& oi = fa ? f.noofa_i : f.noobf_i;
        std::vector<int64_t>& ov = fa ? f.noofa_v : f.noobf_v;
        if (!r.read(data, cnt.data(), cnt.size(), swap)) return bad("cannot read node counts");
        oi.assign(size_t(nf) + 1, 0);
        for (int64_t i = 0; i < nf; ++i) {
          if (cnt[size_t(i)] < 0) return bad("negative node count");
          oi[size_t(i) + 1] = oi[size_t(i)] + cnt[size_t(i)];
        }
        if (oi.back() != nv) return bad("node counts do not add up");
        ov.resize(size_t(nv));
        if (!r.read(data + 4 * nf, ov.data(), ov.size(), swap)) return bad("cannot read node lists");
        break;
      }
      case kPbiPno:
        if (!have_counts || !counted(f.nno_p) || !need(8 * f.nno_p)) return bad("bad pbi_pno record");
        f.pbi_pno.resize(size_t(f.nno_p));
        if (!r.read(data, f.pbi_pno.data(), f.pbi_pno.size(), swap)) return bad("cannot read pbi_pno");
        break;
      case kXNo:
        if (!have_counts || !counted(f.nno) || h.idata[2] != 3 || !need(24 * f.nno)) return bad("bad x_no record");
        f.x_no.resize(size_t(3 * f.nno));
        if (!r.read(data, f.x_no.data(), f.x_no.size(), swap)) return bad("cannot read x_no");
        break;
      case kCvD1:
      case kCvD2:
      case kFaD2:
      case kBfD1:
      case kBfD2:
      case kBfDN33: {
        if (!have_counts) return bad("data record before the counts");
        std::vector<double>* dst = nullptr;
        int64_t rows = 0, width = 1;
        if (h.id == kCvD1 || h.id == kCvD2) {
          rows = f.ncv;
          width = h.id == kCvD1 ? 1 : 3;
          if (h.id == kCvD1 && (name == "vol_cv" || name == "VOL_VV" || name == "vol_vv")) dst = &f.vol_cv;
          if (h.id == kCvD1 && name == "r_vv") dst = &f.r_vv;
          if (h.id == kCvD2 && (name == "x_vv" || name == "X_VV")) dst = &f.x_vv;
          if (h.id == kCvD2 && name == "x_cv") dst = &f.x_cv;
        } else if (h.id == kFaD2) {
          rows = f.nfa;
          width = 3;
          if (name == "n_fa") dst = &f.n_fa;
          if (name == "x_fa") dst = &f.x_fa;
        } else {
          rows = f.nbf;
          width = h.id == kBfD1 ? 1 : (h.id == kBfD2 ? 3 : 9);
          if (h.id == kBfD1 && name == "area_bf") dst = &f.area_bf;
          if (h.id == kBfD1 && name == "area_over_delta_bf") dst = &f.area_over_delta_bf;
          if (h.id == kBfD2 && name == "n_bf") dst = &f.n_bf;
          if (h.id == kBfD2 && name == "x_bf") dst = &f.x_bf;
          if (h.id == kBfDN33 && name == "Gij_bf") dst = &f.Gij_bf;
        }
        if (!dst) break;  // a record of a kind this solver does not use
        if (!counted(rows) || (width == 3 && h.idata[2] != 3) || (width == 9 && (h.idata[2] != 3 || h.idata[3] != 3)) ||
            !need(8 * rows * width))
          return bad("bad record \"" + name + "\"");
        dst->resize(size_t(rows * width));
        if (!r.read(data, dst->data(), dst->size(), swap)) return bad("cannot read \"" + name + "\"");
        break;
      }
      case kEof:
        done = true;
        break;
      default:
        break;  // records this solver does not use (timestamp, params, hashes, surface, ...)
    }
    if (!done) off += h.skip;
  }
  if (!have_counts) return bad("no counts record");
  if (nbf_check != f.nbf) return bad("boundary zone sizes do not add up to nbf");
  if (!f.cv_zones.empty() && ncv_check != f.ncv) return bad("cell zone sizes do not add up to ncv");
  for (int64_t v : f.cvofa)
    if (v < 0 || v >= f.ncv) return bad("cvofa index out of range");
  for (int64_t v : f.cvobf)
    if (v < 0 || v >= f.ncv) return bad("cvobf index out of range");
  for (int64_t v : f.noofa_v)
    if (v < 0 || v >= f.nno) return bad("noofa index out of range");
  for (int64_t v : f.noobf_v)
    if (v < 0 || v >= f.nno) return bad("noobf index out of range");
  return "";
}

std::string voronoi_from_mles(const Mles& f, Voronoi& m) {
  if (f.ncv <= 0 || f.nfa <= 0) return "empty mesh";
  if (f.ncv > 2147483647LL / 2 || f.nfa > 2147483647LL / 3) return "mesh too large for 32-bit cell indices";
  if (f.x_cv.empty() || f.vol_cv.empty() || f.cvofa.empty() || f.n_fa.empty() || f.x_fa.empty())
    return "mesh file lacks one of x_cv, vol_cv, cvofa, n_fa, x_fa";
  // Cartesian transforms along the coordinate axes: three for a periodic box,
  // fewer (or none) when boundary faces close the other directions
  if (f.nbf == 0 && f.npt != 3) return "need three periodic transforms (a box periodic in x, y and z) or boundary faces";
  int axis[3] = {0, 1, 2}, sgn[3] = {1, 1, 1};
  double L[3] = {0, 0, 0};
  for (int i = 0; i < f.npt; ++i) {
    if (f.pt_kind[i] != kTransformCart) return "only Cartesian periodic transforms are supported";
    int nz = 0;
    for (int d = 0; d < 3; ++d)
      if (f.pt_data[i][d] != 0.0) {
        axis[i] = d;
        ++nz;
      }
    if (nz != 1) return "periodic translations must be along the coordinate axes";
    const double t = f.pt_data[i][axis[i]];
    sgn[i] = t > 0 ? 1 : -1;
    if (L[axis[i]] != 0.0) return "two periodic translations along one axis";
    L[axis[i]] = std::fabs(t);
  }
  const int ncv = int(f.ncv), nfa = int(f.nfa);
  std::vector<int> cvofa(static_cast<size_t>(2 * nfa)), shift(static_cast<size_t>(3 * nfa), 0);
  int64_t zone_end = 0;
  int zone = -1;
  for (int fa = 0; fa < nfa; ++fa) {
    while (fa >= zone_end) {
      ++zone;
      zone_end += f.nfa_zone[zone];
    }
    cvofa[size_t(2 * fa)] = int(f.cvofa[size_t(2 * fa)]);
    cvofa[size_t(2 * fa + 1)] = int(f.cvofa[size_t(2 * fa + 1)]);
    if (zone < 26) {
      const int bits = kZoneBits[zone];
      for (int i = 0; i < 3; ++i) {
        if (bits & (1 << (2 * i))) shift[size_t(3 * fa + axis[i])] += sgn[i];
        if (bits & (1 << (2 * i + 1))) shift[size_t(3 * fa + axis[i])] -= sgn[i];
      }
    }
  }
  std::string e = build_from_arrays(m, ncv, f.x_cv.data(), f.vol_cv.data(), nfa, cvofa.data(), shift.data(),
                                    f.n_fa.data(), f.x_fa.data(), L);
  if (!e.empty()) return e;
  if (f.x_vv.size() == size_t(3 * ncv)) m.x_vv = f.x_vv;
  if (f.r_vv.size() == size_t(ncv)) m.r_vv = f.r_vv;
  if (f.nbf > 0) {
    // boundary zones in file order, each a contiguous range of boundary faces
    const int nbf = int(f.nbf);
    if (f.cvobf.size() != size_t(nbf) || f.n_bf.size() != size_t(3 * nbf) || f.x_bf.size() != size_t(3 * nbf))
      return "mesh file lacks one of cvobf, n_bf, x_bf";
    if (f.x_vv.size() != size_t(3 * ncv) &&
        (f.area_over_delta_bf.size() != size_t(nbf) || f.Gij_bf.size() != size_t(9 * nbf)))
      return "mesh file lacks x_vv to build its boundary-face operators";
    m.nbf = nbf;
    m.bf_zone_i.assign(1, 0);
    for (const auto& z : f.bf_zones) {
      if (z.ibf_f != m.bf_zone_i.back()) return "boundary zones are not contiguous in face order";
      m.bf_zone_name.push_back(z.name);
      m.bf_zone_i.push_back(int(z.ibf_f + z.nbf));
    }
    if (m.bf_zone_i.back() != nbf) return "boundary zone counts do not add up to nbf";
    m.cvobf.resize(size_t(nbf));
    for (int b = 0; b < nbf; ++b) {
      m.cvobf[size_t(b)] = int(f.cvobf[size_t(b)]);
      if (m.cvobf[size_t(b)] < 0 || m.cvobf[size_t(b)] >= ncv) return "cvobf index out of range";
    }
    m.n_bf = f.n_bf;
    m.x_bf = f.x_bf;
    if (f.area_over_delta_bf.size() == size_t(nbf)) m.aod_bf_file = f.area_over_delta_bf;
    if (f.Gij_bf.size() == size_t(9 * nbf)) m.gij_bf_file = f.Gij_bf;
  }
  // the domain extents over the node records (Sponge::initData,
  // IdealGasSolverBcs.cpp:4104-4113), when the file holds them
  if (!f.x_no.empty()) {
    double mx[5] = {-1.0E-20, -1.0E-20, -1.0E-20, -1.0E-20, -1.0E-20};
    size_t arg[5] = {0, 0, 0, 0, 0};
    for (size_t i = 0; i + 2 < f.x_no.size(); i += 3) {
      const double* x = &f.x_no[i];
      const double v[5] = {x[0], x[1], x[2], std::sqrt(x[1] * x[1] + x[2] * x[2]),
                           std::sqrt(x[0] * x[0] + x[1] * x[1])};
      for (int k = 0; k < 5; ++k)
        if (mx[k] < v[k]) {
          mx[k] = v[k];
          arg[k] = i;
        }
    }
    for (int k = 0; k < 5; ++k) m.x_max[k] = mx[k];
    m.has_x_max = true;
    m.x_ext_nodes.clear();
    for (int k = 0; k < 5; ++k) {
      bool seen = false;
      for (int j = 0; j < k; ++j) seen = seen || arg[j] == arg[k];
      if (!seen) m.x_ext_nodes.insert(m.x_ext_nodes.end(), &f.x_no[arg[k]], &f.x_no[arg[k]] + 3);
    }
  }
  return "";
}

namespace {

class Writer {
 public:
  explicit Writer(const std::string& path) { fp_ = std::fopen(path.c_str(), "wb"); }
  ~Writer() {
    if (fp_) std::fclose(fp_);
  }
  bool ok() const { return fp_ != nullptr && good_; }
  bool close() {
    const bool r = fp_ && good_ && std::fclose(fp_) == 0;
    fp_ = nullptr;
    return r;
  }
  void start(int32_t magic, int32_t version) {
    const int32_t h[2] = {magic, version};
    put(h, 2);
  }
  // a record header followed by `bytes` of data the caller writes next
  RecHeader header(const char* name, int32_t id, int64_t bytes) {
    RecHeader h;
    std::memset(&h, 0, sizeof(h));
    std::snprintf(h.name, sizeof(h.name), "%s", name);
    h.id = id;
    h.skip = 256 + bytes;
    return h;
  }
  void put_header(const RecHeader& h) { put(&h, 1); }
  template <class T>
  void put(const T* p, size_t n) {
    if (fp_ && good_ && n) good_ = std::fwrite(p, sizeof(T), n, fp_) == n;
  }
  // per-cell (or per-face) array of width 1 or 3
  void put_array(const char* name, int32_t id, const double* p, int64_t rows, int width) {
    RecHeader h = header(name, id, 8 * rows * width);
    set_count(h.idata, rows);
    if (width == 3) h.idata[2] = 3;
    if (width == 9) h.idata[2] = h.idata[3] = 3;
    put_header(h);
    put(p, size_t(rows * width));
  }

 private:
  FILE* fp_ = nullptr;
  bool good_ = true;
};

}  // namespace

std::string write_mles(const Voronoi& m, const std::string& path) {
  const int ncv = m.ncv, nfa = m.nfa;
  if (ncv <= 0 || nfa <= 0) return "empty mesh";
  // periodic bits of a face: transform d translates by +L[d] along axis d
  int zone_of_bits[64];
  for (int& z : zone_of_bits) z = -1;
  zone_of_bits[0] = 26;
  for (int z = 0; z < 26; ++z) zone_of_bits[kZoneBits[z]] = z;
  std::vector<int> zone(static_cast<size_t>(nfa));
  int64_t nfa_zone[27] = {0}, noofa_zone[27] = {0};
  for (int fa = 0; fa < nfa; ++fa) {
    int bits = 0;
    for (int d = 0; d < 3; ++d) {
      const int s = m.shift[size_t(3 * fa + d)];
      if (s == 1) bits |= 1 << (2 * d);
      if (s == -1) bits |= 1 << (2 * d + 1);
    }
    zone[size_t(fa)] = zone_of_bits[bits];
    if (zone[size_t(fa)] < 0) return "face with an unsupported periodic shift";
    ++nfa_zone[zone[size_t(fa)]];
  }
  // faces in zone order (periodic zones first, interior faces last), stable within a zone
  std::vector<int> order(static_cast<size_t>(nfa));
  for (int fa = 0; fa < nfa; ++fa) order[size_t(fa)] = fa;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return zone[size_t(a)] < zone[size_t(b)]; });
  std::vector<int64_t> cvofa(static_cast<size_t>(2 * nfa));
  std::vector<double> n_fa(static_cast<size_t>(3 * nfa)), x_fa(static_cast<size_t>(3 * nfa));
  for (int i = 0; i < nfa; ++i) {
    const int fa = order[size_t(i)];
    cvofa[size_t(2 * i)] = m.cvofa[size_t(2 * fa)];
    cvofa[size_t(2 * i + 1)] = m.cvofa[size_t(2 * fa + 1)];
    for (int d = 0; d < 3; ++d) {
      n_fa[size_t(3 * i + d)] = m.n_fa[size_t(3 * fa + d)];
      x_fa[size_t(3 * i + d)] = m.x_fa[size_t(3 * fa + d)];
    }
  }

  bool periodic = false;
  for (int8_t sh : m.shift) periodic = periodic || sh != 0;
  const int nbf = m.nbf;
  if (nbf > 0 && (m.cvobf.size() != size_t(nbf) || m.n_bf.size() != size_t(3 * nbf) ||
                  m.x_bf.size() != size_t(3 * nbf) || m.bf_zone_i.size() != m.bf_zone_name.size() + 1))
    return "inconsistent boundary faces";

  // node records: only the nodes that attain the domain's extents
  const int64_t nno = m.has_x_max ? int64_t(m.x_ext_nodes.size() / 3) : 0;

  Writer w(path);
  if (!w.ok()) return "cannot create " + path;
  w.start(kMagic, 5);
  {
    RecHeader h = w.header("no_fa_bf_cv_counts", kCounts, 2 * 27 * 8);
    set_count(h.idata + 0, nno);  // nno: the extents' nodes, no connectivity
    set_count(h.idata + 2, nbf);
    set_count(h.idata + 4, nfa);
    set_count(h.idata + 6, ncv);
    set_count(h.idata + 8, 0);   // nno_p
    set_count(h.idata + 10, 0);  // nno_pb
    w.put_header(h);
    w.put(nfa_zone, 27);
    w.put(noofa_zone, 27);
  }
  if (periodic) {
    RecHeader h = w.header("periodic_transforms", kPeriodic, 0);
    h.idata[0] = 3;
    for (int d = 0; d < 3; ++d) {
      h.idata[1 + d] = kTransformCart;
      h.rdata[3 * d + d] = m.L[d];
    }
    w.put_header(h);
  }
  // boundary zones (a contiguous range of boundary faces each) and their faces
  for (size_t z = 0; z < m.bf_zone_name.size(); ++z) {
    RecHeader h = w.header(m.bf_zone_name[z].c_str(), kBfZoneHeader, 0);
    h.idata[1] = int32_t(z);
    set_count(h.idata + 2, m.bf_zone_i[z + 1] - m.bf_zone_i[z]);
    set_count(h.idata + 4, 0);  // no node records
    set_count(h.idata + 6, m.bf_zone_i[z]);
    w.put_header(h);
  }
  if (nbf > 0) {
    std::vector<int64_t> cvobf(m.cvobf.begin(), m.cvobf.end());
    RecHeader h = w.header("cvobf_global", kCvobf, 8 * int64_t(nbf));
    set_count(h.idata, nbf);
    w.put_header(h);
    w.put(cvobf.data(), cvobf.size());
    w.put_array("n_bf", kBfD2, m.n_bf.data(), nbf, 3);
    w.put_array("x_bf", kBfD2, m.x_bf.data(), nbf, 3);
    if (m.aod_bf_file.size() == size_t(nbf))
      w.put_array("area_over_delta_bf", kBfD1, m.aod_bf_file.data(), nbf, 1);
    if (m.gij_bf_file.size() == size_t(9 * nbf)) w.put_array("Gij_bf", kBfDN33, m.gij_bf_file.data(), nbf, 9);
  }
  {
    RecHeader h = w.header("cvofa_global", kCvofa, 16 * int64_t(nfa));
    set_count(h.idata, nfa);
    h.idata[2] = 2;
    w.put_header(h);
    w.put(cvofa.data(), cvofa.size());
  }
  {
    RecHeader h = w.header("fluid", kCvZoneHeader, 0);
    h.idata[0] = kZoneFluid;
    h.idata[1] = 0;
    set_count(h.idata + 2, ncv);
    set_count(h.idata + 4, 0);
    w.put_header(h);
  }
  if (nno > 0) w.put_array("x_no", kXNo, m.x_ext_nodes.data(), nno, 3);
  w.put_array("vol_cv", kCvD1, m.vol.data(), ncv, 1);
  w.put_array("x_vv", kCvD2, m.x_vv.data(), ncv, 3);
  if (m.r_vv.size() == size_t(ncv)) w.put_array("r_vv", kCvD1, m.r_vv.data(), ncv, 1);
  w.put_array("x_cv", kCvD2, m.x_cv.data(), ncv, 3);
  w.put_array("n_fa", kFaD2, n_fa.data(), nfa, 3);
  w.put_array("x_fa", kFaD2, x_fa.data(), nfa, 3);
  w.put_header(w.header("EOF", kEof, 0));
  if (!w.close()) return "write failed: " + path;
  return "";
}

std::string write_sles(const std::string& path, const SlesData& s) {
  const int64_t ncv = s.ncv;
  if (ncv <= 0) return "empty state";
  Writer w(path);
  if (!w.ok()) return "cannot create " + path;
  // data files carry the magic number plus one (StaticSolver.cpp:7575-7577)
  w.start(kMagic + 1, 5);
  {
    RecHeader h = w.header("step", kI0, 0);
    h.idata[0] = s.step;
    w.put_header(h);
  }
  {
    RecHeader h = w.header("time", kD0, 0);
    h.rdata[0] = s.time;
    w.put_header(h);
  }
  {
    RecHeader h = w.header("dt", kD0, 0);
    h.rdata[0] = s.dt;
    w.put_header(h);
  }
  w.put_array("rho", kCvD1, s.rho, ncv, 1);
  w.put_array("u", kCvD2, s.u, ncv, 3);
  w.put_array("rhoE", kCvD1, s.rhoE, ncv, 1);
  if (s.p) w.put_array("p", kCvD1, s.p, ncv, 1);
  if (s.T) w.put_array("T", kCvD1, s.T, ncv, 1);
  if (s.lijmij) w.put_array("lijmij", kCvD1, s.lijmij, ncv, 1);
  if (s.mijmij) w.put_array("mijmij", kCvD1, s.mijmij, ncv, 1);
  // Sponge's registered cv data (IdealGasSolverBcs.cpp:4089-4090)
  if (s.T_sponge) w.put_array("T_sponge", kCvD1, s.T_sponge, ncv, 1);
  if (s.u_sponge) w.put_array("u_sponge", kCvD2, s.u_sponge, ncv, 3);
  for (size_t i = 0; i < s.sc_names.size() && i < s.sc.size(); ++i)
    w.put_array(s.sc_names[i].c_str(), kCvD1, s.sc[i], ncv, 1);
  w.put_header(w.header("EOF", kEof, 0));
  if (!w.close()) return "write failed: " + path;
  return "";
}

std::string read_sles(const std::string& path, int64_t ncv, SlesRead& s) {
  s = SlesRead();
  Reader r(path);
  if (!r.ok()) return "cannot open " + path;
  const int64_t fsize = r.size();
  int32_t head[2] = {0, 0};
  if (!r.read(0, head, 2, false)) return path + ": too short for a record file";
  bool swap = false;
  if (head[0] != kMagic + 1) {
    swap_bytes(head, 2);
    if (head[0] != kMagic + 1) return path + ": not a data record file (magic number)";
    swap = true;
  }
  if (head[1] != 5) return path + ": unknown io version " + std::to_string(head[1]);
  auto bad = [&](const std::string& what) { return path + ": " + what; };
  int64_t off = 8;
  for (;;) {
    RecHeader h;
    if (off + 256 > fsize || !r.read(off, &h, 1, false)) return bad("truncated (no EOF record)");
    if (swap) swap_header(h);
    if (h.id == kEof) break;
    if (h.skip < 256 || off + h.skip > fsize) return bad("record runs past the end of the file");
    const std::string name = hname(h);
    if (h.id == kI0 && name == "step") {
      s.step = h.idata[0];
      s.have_step = true;
    } else if (h.id == kD0 && name == "time") {
      s.time = h.rdata[0];
      s.have_time = true;
    } else if (h.id == kD0 && name == "dt") {
      s.dt = h.rdata[0];
      s.have_dt = true;
    } else if (h.id == kCvD1 || h.id == kCvD2) {
      std::vector<double>* dst = nullptr;
      if (h.id == kCvD1 && name == "rho") dst = &s.rho;
      if (h.id == kCvD2 && name == "u") dst = &s.u;
      if (h.id == kCvD1 && name == "rhoE") dst = &s.rhoE;
      if (h.id == kCvD1 && name == "p") dst = &s.p;
      if (h.id == kCvD1 && name == "T") dst = &s.T;
      if (h.id == kCvD1 && name == "lijmij") dst = &s.lijmij;
      if (h.id == kCvD1 && name == "mijmij") dst = &s.mijmij;
      if (h.id == kCvD1 && name == "T_sponge") dst = &s.T_sponge;
      if (h.id == kCvD2 && name == "u_sponge") dst = &s.u_sponge;
      if (h.id == kCvD1 && !dst) {
        s.other.emplace_back(name, std::vector<double>());
        dst = &s.other.back().second;
      }
      if (dst) {
        const int width = h.id == kCvD1 ? 1 : 3;
        int64_t n = 0;
        if (!get_count(h.idata, n) || (width == 3 && h.idata[2] != 3) || 8 * n * width > h.skip - 256)
          return bad("bad record \"" + name + "\"");
        if (n != ncv)
          return bad("\"" + name + "\" holds " + std::to_string(n) + " cells, the mesh has " + std::to_string(ncv));
        dst->resize(size_t(n * width));
        if (!r.read(off + 256, dst->data(), dst->size(), swap)) return bad("cannot read \"" + name + "\"");
      }
    }
    off += h.skip;
  }
  if (s.rho.empty() || s.u.empty() || s.rhoE.empty()) return bad("lacks one of rho, u, rhoE");
  return "";
}

}  // namespace cgmesh
