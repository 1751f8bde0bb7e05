// Setup products on disk.  build_setup is deterministic in the mesh (the
// connectivity, and the initial positions for the BVH split) and in the layout
// constants, so its result can be stored once per mesh and loaded by every
// later cloth of the same mesh: a serving process admits a request of a known
// mesh without rebuilding its pattern, level schedules, sweep slots, gather
// lists and BVH.
//
// File: "<dir>/<key>.cssetup", key = two independent 64-bit FNV-1a hashes of
// the inputs and the layout constants.  Contents: a header (magic, version,
// sizes, both hashes), then every MeshSetup field in a fixed order (scalars
// as int64, vectors as an int64 count followed by the elements), then a
// 64-bit FNV-1a hash of those field bytes.  A file that is missing,
// truncated, for another mesh, whose payload hash differs, or whose contents
// fail validate_setup (indices out of range, broken schedules) is ignored
// and rebuilt: nothing read from disk reaches a kernel unchecked.  Writes go
// to a temporary name first and are renamed into place.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "cs_setup.hpp"

namespace cs {

namespace {

constexpr char kMagic[8] = {'C', 'S', 'S', 'E', 'T', 'U', 'P', '1'};
constexpr long long kVersion = 6;  // bump when MeshSetup's fields or their meaning change

struct Fnv {
  unsigned long long h;
  explicit Fnv(unsigned long long seed) : h(seed) {}
  void bytes(const void* p, size_t n) {
    const unsigned char* b = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < n; i++) {
      h ^= b[i];
      h *= 1099511628211ull;
    }
  }
  template <class T>
  void value(T v) { bytes(&v, sizeof(v)); }
};

void hash_inputs(Fnv& f, int np, int nt, int nq, const int* tri, const int* quad,
                 const double* xyz) {
  f.value(kVersion);
  f.value(kRingRows);
  f.value(kQuadRows);
  f.value(kMaxBlocksPerRow);
  f.value(np);
  f.value(nt);
  f.value(nq);
  f.bytes(tri, sizeof(int) * 3 * (size_t)nt);
  if (nq > 0) f.bytes(quad, sizeof(int) * 4 * (size_t)nq);
  f.bytes(xyz, sizeof(double) * 3 * (size_t)np);
}

// one list of the fields for both directions
template <class IO>
bool fields(IO& io, MeshSetup& s) {
  return io.num(s.np) && io.num(s.nt) && io.num(s.nq) && io.num(s.nnz) &&
         io.vec(s.rowptr) && io.vec(s.colind) && io.vec(s.diaind) && io.vec(s.rowof) &&
         io.vec(s.lf_ptr) && io.vec(s.lf_rows) && io.vec(s.lb_ptr) && io.vec(s.lb_rows) &&
         io.num(s.nlf) && io.num(s.nlb) && io.num(s.max_lf) && io.num(s.max_lb) &&
         io.vec(s.fu_ptr) && io.vec(s.fu_kj) && io.vec(s.fu_code) &&
         io.num(s.ncomp) && io.vec(s.comp) && io.vec(s.sf_rows) && io.vec(s.sb_rows) &&
         io.vec(s.qf_ptr) && io.vec(s.qb_ptr) && io.vec(s.qf_comp) && io.vec(s.qb_comp) &&
         io.num(s.nqf) && io.num(s.nqb) && io.num(s.qwidth) &&
         io.num(s.kl) && io.num(s.ku) && io.vec(s.f_rci) && io.vec(s.b_rc) && io.vec(s.fslot) &&
         io.vec(s.f_meta) && io.vec(s.b_meta) && io.flag(s.ring_ok) &&
         io.vec(s.crs_src_ptr) && io.vec(s.crs_src) && io.vec(s.dia_src_ptr) &&
         io.vec(s.dia_src) && io.vec(s.grad_src_ptr) && io.vec(s.grad_src) &&
         io.vec(s.ring_ptr) && io.vec(s.ring) && io.num(s.nnode) && io.num(s.root) &&
         io.vec(s.node) && io.vec(s.pre) && io.vec(s.pre_end) && io.vec(s.by_pre) &&
         io.vec(s.leaf_of) && io.vec(s.depth) && io.vec(s.refit_order) &&
         io.vec(s.refit_lvl_ptr) && io.num(s.max_depth);
}

struct Writer {
  FILE* f;
  Fnv payload{0xcbf29ce484222325ull};
  bool raw(const void* p, size_t n) {
    payload.bytes(p, n);
    return n == 0 || fwrite(p, 1, n, f) == n;
  }
  bool num(int v) {
    const long long x = v;
    return raw(&x, sizeof(x));
  }
  bool flag(bool v) { return num(v ? 1 : 0); }
  template <class T>
  bool vec(const std::vector<T>& v) {
    const long long n = (long long)v.size();
    return raw(&n, sizeof(n)) && raw(v.data(), sizeof(T) * v.size());
  }
};

struct Reader {
  FILE* f;
  long long left;  // bytes left in the file
  Fnv payload{0xcbf29ce484222325ull};
  bool raw(void* p, size_t n) {
    if ((long long)n > left) return false;
    left -= (long long)n;
    if (n != 0 && fread(p, 1, n, f) != n) return false;
    payload.bytes(p, n);
    return true;
  }
  bool num(int& v) {
    long long x = 0;
    if (!raw(&x, sizeof(x)) || x < -2147483648LL || x > 2147483647LL) return false;
    v = (int)x;
    return true;
  }
  bool flag(bool& v) {
    int x = 0;
    if (!num(x)) return false;
    v = x != 0;
    return true;
  }
  template <class T>
  bool vec(std::vector<T>& v) {
    long long n = 0;
    if (!raw(&n, sizeof(n)) || n < 0 || n > left / (long long)sizeof(T)) return false;
    v.resize((size_t)n);
    return raw(v.data(), sizeof(T) * (size_t)n);
  }
};

struct Header {
  char magic[8];
  long long version, np, nt, nq;
  unsigned long long h1, h2;
};

std::mutex g_mu;
std::string g_dir;
long long g_hits = 0, g_misses = 0, g_writes = 0;

// build_setup, then validate_setup when CLOTHSIM_VALIDATE_SETUP is set to a
// value other than "0" (a debug check of what the kernels rely on)
std::string build_setup_checked(MeshSetup& s, int np, int nt, int nq, const int* tri,
                                const int* quad, const double* xyz) {
  std::string err = build_setup(s, np, nt, nq, tri, quad, xyz);
  if (!err.empty()) return err;
  const char* v = std::getenv("CLOTHSIM_VALIDATE_SETUP");
  if (v && *v && std::strcmp(v, "0") != 0) {
    err = validate_setup(s, np, nt, nq, tri, quad);
    if (!err.empty()) return "setup failed validation: " + err;
  }
  return std::string();
}

}  // namespace

void set_setup_cache_dir(const char* dir) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_dir = dir ? dir : "";
}

void setup_cache_stats(long long* hits, long long* misses, long long* writes) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (hits) *hits = g_hits;
  if (misses) *misses = g_misses;
  if (writes) *writes = g_writes;
}

std::string build_setup_cached(MeshSetup& s, int np, int nt, int nq, const int* tri,
                               const int* quad, const double* xyz) {
  std::string dir;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    dir = g_dir;
  }
  if (dir.empty() || np <= 0 || nt <= 0 || !tri || !xyz || (nq > 0 && !quad))
    return build_setup_checked(s, np, nt, nq, tri, quad, xyz);
  Fnv a(1469598103934665603ull), b(0x9e3779b97f4a7c15ull);
  hash_inputs(a, np, nt, nq, tri, quad, xyz);
  hash_inputs(b, np, nt, nq, tri, quad, xyz);
  char name[64];
  snprintf(name, sizeof(name), "%016llx%016llx.cssetup", a.h, b.h);
  const std::string path = dir + "/" + name;
  if (FILE* f = fopen(path.c_str(), "rb")) {
    fseek(f, 0, SEEK_END);
    const long long size = ftell(f);
    fseek(f, 0, SEEK_SET);
    Reader r{f, size};
    Header h;
    MeshSetup t;
    bool ok = r.raw(&h, sizeof(h)) && std::memcmp(h.magic, kMagic, 8) == 0 &&
              h.version == kVersion && h.np == np && h.nt == nt && h.nq == nq &&
              h.h1 == a.h && h.h2 == b.h;
    if (ok) {
      r.payload = Fnv(0xcbf29ce484222325ull);
      ok = fields(r, t);
      const unsigned long long want = r.payload.h;
      unsigned long long got = 0;
      ok = ok && r.raw(&got, sizeof(got)) && r.left == 0 && got == want;
    }
    fclose(f);
    ok = ok && validate_setup(t, np, nt, nq, tri, quad).empty();
    if (ok) {
      s = std::move(t);
      std::lock_guard<std::mutex> lk(g_mu);
      g_hits++;
      return std::string();
    }
  }
  {
    std::lock_guard<std::mutex> lk(g_mu);
    g_misses++;
  }
  std::string err = build_setup_checked(s, np, nt, nq, tri, quad, xyz);
  if (!err.empty()) return err;
  // best effort: a cache that cannot be written only costs the rebuild
  char tmp_name[96];
  snprintf(tmp_name, sizeof(tmp_name), "%s.tmp%p", name, (void*)&s);
  const std::string tmp = dir + "/" + tmp_name;
  if (FILE* f = fopen(tmp.c_str(), "wb")) {
    Writer w{f};
    Header h;
    std::memcpy(h.magic, kMagic, 8);
    h.version = kVersion;
    h.np = np;
    h.nt = nt;
    h.nq = nq;
    h.h1 = a.h;
    h.h2 = b.h;
    bool ok = w.raw(&h, sizeof(h));
    w.payload = Fnv(0xcbf29ce484222325ull);
    ok = ok && fields(w, s);
    const unsigned long long sum = w.payload.h;
    ok = ok && w.raw(&sum, sizeof(sum));
    const bool closed = fclose(f) == 0;
    if (ok && closed && std::rename(tmp.c_str(), path.c_str()) == 0) {
      std::lock_guard<std::mutex> lk(g_mu);
      g_writes++;
    } else {
      std::remove(tmp.c_str());
    }
  }
  return std::string();
}

}  // namespace cs
