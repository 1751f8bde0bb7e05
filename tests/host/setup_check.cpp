// Host-only check of the per-mesh setup and its cache (run under ASan/UBSan):
//   setup_check <mesh.bin> <cache dir>
// mesh.bin: int32 np, nt, nq; double xyz[3np]; int32 tri[3nt], quad[4nq].
// 1. build_setup, then validate_setup must accept it;
// 2. planted mistakes in copies of the setup (an index out of range, a ring
//    offset past the LDS ring, a level order that breaks a dependency, a BVH
//    child out of range, a gather source past the elements) must each be
//    rejected;
// 3. through the cache: a miss writes the file, a second lookup hits and
//    returns the same setup, a file with one flipped payload byte is
//    rejected and rebuilt (a miss), and the rebuilt setup equals the fresh one.
// Prints "ok" and exits 0 when every check holds.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "cs_setup.hpp"

using cs::MeshSetup;

static int fail(const std::string& m) {
  std::fprintf(stderr, "setup_check: %s\n", m.c_str());
  return 1;
}

static bool same(const MeshSetup& a, const MeshSetup& b) {
  return a.np == b.np && a.nnz == b.nnz && a.rowptr == b.rowptr && a.colind == b.colind &&
         a.diaind == b.diaind && a.lf_rows == b.lf_rows && a.lb_rows == b.lb_rows &&
         a.fu_kj == b.fu_kj && a.fu_code == b.fu_code && a.sf_rows == b.sf_rows &&
         a.qf_ptr == b.qf_ptr && a.f_meta == b.f_meta && a.b_meta == b.b_meta &&
         a.ring_ok == b.ring_ok && a.crs_src == b.crs_src && a.dia_src == b.dia_src &&
         a.grad_src == b.grad_src && a.ring == b.ring && a.node == b.node && a.pre == b.pre &&
         a.refit_order == b.refit_order;
}

int main(int argc, char** argv) {
  if (argc != 3) return fail("usage: setup_check mesh.bin cachedir");
  FILE* f = std::fopen(argv[1], "rb");
  if (!f) return fail("cannot open mesh");
  int hdr[3];
  if (std::fread(hdr, sizeof(int), 3, f) != 3) return fail("short mesh header");
  const int np = hdr[0], nt = hdr[1], nq = hdr[2];
  std::vector<double> xyz(3 * (size_t)np);
  std::vector<int> tri(3 * (size_t)nt), quad(4 * (size_t)nq);
  bool rd = std::fread(xyz.data(), sizeof(double), xyz.size(), f) == xyz.size() &&
            std::fread(tri.data(), sizeof(int), tri.size(), f) == tri.size() &&
            std::fread(quad.data(), sizeof(int), quad.size(), f) == quad.size();
  std::fclose(f);
  if (!rd) return fail("short mesh file");

  MeshSetup s;
  std::string err = cs::build_setup(s, np, nt, nq, tri.data(), quad.data(), xyz.data());
  if (!err.empty()) return fail("build_setup: " + err);
  err = cs::validate_setup(s, np, nt, nq, tri.data(), quad.data());
  if (!err.empty()) return fail("validate_setup rejects a fresh setup: " + err);

  // planted mistakes, each in its own copy
  struct Plant {
    const char* what;
    void (*apply)(MeshSetup&);
  };
  const Plant plants[] = {
      {"column index out of range", [](MeshSetup& t) { t.colind[t.nnz / 2] = t.np; }},
      {"ring offset past the ring",
       [](MeshSetup& t) { if (t.kl > 0) t.f_meta[0] = 0xfff8; else t.colind[0] = -1; }},
      {"level order breaking a dependency",
       [](MeshSetup& t) {
         // move the last row of the forward schedule into level 0
         const int last = t.lf_rows.back();
         t.lf_rows.pop_back();
         t.lf_rows.insert(t.lf_rows.begin(), last);
         for (int l = 1; l < t.nlf; l++) t.lf_ptr[l]++;
       }},
      {"BVH child out of range", [](MeshSetup& t) { t.node[2] = t.nnode; }},
      {"gather source past the elements", [](MeshSetup& t) { t.dia_src[0] = 9 * t.nt; }},
      {"vertex ring entry out of range", [](MeshSetup& t) { t.ring[0] = t.np; }},
      {"ILU update outside the pattern", [](MeshSetup& t) {
         if (!t.fu_kj.empty()) t.fu_kj[0] = t.nnz; else t.colind[0] = t.np; }},
  };
  for (const Plant& p : plants) {
    MeshSetup t = s;
    p.apply(t);
    if (cs::validate_setup(t, np, nt, nq, tri.data(), quad.data()).empty())
      return fail(std::string("validate_setup accepts a planted mistake: ") + p.what);
  }

  // the cache
  cs::set_setup_cache_dir(argv[2]);
  long long h0, m0, w0, h1, m1, w1;
  cs::setup_cache_stats(&h0, &m0, &w0);
  MeshSetup a, b, c;
  err = cs::build_setup_cached(a, np, nt, nq, tri.data(), quad.data(), xyz.data());
  if (!err.empty()) return fail("cached build: " + err);
  err = cs::build_setup_cached(b, np, nt, nq, tri.data(), quad.data(), xyz.data());
  if (!err.empty()) return fail("cached load: " + err);
  cs::setup_cache_stats(&h1, &m1, &w1);
  if (h1 != h0 + 1 || m1 != m0 + 1 || w1 != w0 + 1) return fail("expected one miss, one write, one hit");
  if (!same(a, s) || !same(b, s)) return fail("cached setup differs from the fresh one");
  // flip one payload byte of the file (the only .cssetup in the directory)
  std::string path;
  {
    std::string cmd = std::string("ls ") + argv[2] + "/*.cssetup";
    FILE* p = popen(cmd.c_str(), "r");
    char buf[4096];
    if (p && std::fgets(buf, sizeof(buf), p)) {
      path = buf;
      while (!path.empty() && (path.back() == '\n' || path.back() == '\r')) path.pop_back();
    }
    if (p) pclose(p);
  }
  if (path.empty()) return fail("cache file not found");
  {
    FILE* g = std::fopen(path.c_str(), "r+b");
    if (!g) return fail("cannot open cache file");
    std::fseek(g, 0, SEEK_END);
    const long size = std::ftell(g);
    const long at = size / 2;
    std::fseek(g, at, SEEK_SET);
    int ch = std::fgetc(g);
    std::fseek(g, at, SEEK_SET);
    std::fputc(ch ^ 0x10, g);
    std::fclose(g);
  }
  err = cs::build_setup_cached(c, np, nt, nq, tri.data(), quad.data(), xyz.data());
  if (!err.empty()) return fail("rebuild after corruption: " + err);
  long long h2, m2, w2;
  cs::setup_cache_stats(&h2, &m2, &w2);
  if (h2 != h1 || m2 != m1 + 1) return fail("a corrupted cache file was accepted");
  if (!same(c, s)) return fail("rebuilt setup differs from the fresh one");
  cs::set_setup_cache_dir(nullptr);
  std::printf("ok np=%d nt=%d nq=%d nnz=%d levels=%d/%d ring_ok=%d\n", np, nt, nq, s.nnz, s.nlf,
              s.nlb, s.ring_ok ? 1 : 0);
  return 0;
}
