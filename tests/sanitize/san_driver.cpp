// Host-code sanitizer driver (CPU test infrastructure): built with
// AddressSanitizer and UndefinedBehaviorSanitizer from the product's host
// symbolic analysis (artisynth_core_amd/csrc/symbolic.cpp) and the C oracle
// (oracle/kkt_oracle.c), it runs both on the matrices named on the command line
// and checks them against each other, so every index-heavy host loop of the
// analysis and the oracle runs under the sanitizers on real inputs.
//
//   san_driver IN.bin OUT.bin     one matrix (format below), both orderings
//   san_driver --bad              malformed patterns through validate_pattern
//
// IN.bin: int32 n, int32 nnz, int32 rowOffs[n+1], int32 colIdxs[nnz],
// float64 vals[nnz], uint8 delay[n], float64 b[n] (0-based upper CSR).
// OUT.bin, per ordering (natural, nested dissection): int32 nsuper,
// int32 perm[n], int32 sup_ptr[nsuper+1], int64 nnzL, float64 residual.
// Exit status 0 when every check passes; a sanitizer report aborts the run.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "kkt_oracle.h"
#include "symbolic.h"

namespace {

int fails = 0;

void check(bool ok, const char* what, int ordering) {
  if (!ok) {
    std::fprintf(stderr, "FAIL (ordering %d): %s\n", ordering, what);
    fails++;
  }
}

template <typename T>
bool rd(FILE* f, T* p, size_t cnt) {
  return cnt == 0 || std::fread(p, sizeof(T), cnt, f) == cnt;
}

double norm2(const std::vector<double>& v) {
  double s = 0.0;
  for (double x : v) s += x * x;
  return std::sqrt(s);
}

int run_matrix(const char* inp, const char* outp) {
  FILE* f = std::fopen(inp, "rb");
  if (!f) return 2;
  int n = 0, nnz = 0;
  if (!rd(f, &n, 1) || !rd(f, &nnz, 1) || n <= 0 || nnz <= 0) return 2;
  std::vector<int> ro(n + 1), ci(nnz);
  std::vector<double> va(nnz), b(n);
  std::vector<unsigned char> delay(n);
  if (!rd(f, ro.data(), n + 1) || !rd(f, ci.data(), nnz) || !rd(f, va.data(), nnz) ||
      !rd(f, delay.data(), n) || !rd(f, b.data(), n))
    return 2;
  std::fclose(f);
  FILE* o = std::fopen(outp, "wb");
  if (!o) return 2;
  std::string err;
  check(kkt::validate_pattern(n, nnz, ro.data(), ci.data(), err) == 0, "validate_pattern", -1);
  for (int ordering = 0; ordering < 2; ordering++) {
    kkt::SymOptions opt;
    opt.ordering = ordering;
    kkt::Symbolic S;
    const int rc = kkt::symbolic_analyze(n, ro.data(), ci.data(), delay.data(), opt, S, err);
    check(rc == 0, "symbolic_analyze", ordering);
    if (rc != 0) continue;
    // elimination tree and column counts against the oracle's symbolic elimination
    std::vector<int> par(n), cc(n);
    check(orc_symbolic(n, ro.data(), ci.data(), S.perm.data(), par.data(), cc.data()) == 0,
          "orc_symbolic", ordering);
    check(par == S.parent, "elimination tree", ordering);
    check(cc == S.colcount, "column counts", ordering);
    // supernode row structures against the oracle's union construction
    std::vector<long> rp(S.nsuper + 1);
    std::vector<int> rows(S.rows.size() + 1);
    const long nr = orc_supernode_rows(n, ro.data(), ci.data(), S.perm.data(), S.nsuper, S.sup_ptr.data(),
                                       rp.data(), rows.data(), (long)rows.size());
    check(nr == (long)S.rows.size(), "supernode row count", ordering);
    bool same = nr >= 0;
    for (int s = 0; same && s <= S.nsuper; s++) same = rp[s] == (long)S.rows_ptr[s];
    for (long i = 0; same && i < nr; i++) same = rows[i] == S.rows[i];
    check(same, "supernode rows", ordering);
    // factor, solve and refine in the oracle on the product's partition
    double resid = -1.0;
    orc_factor* F = orc_factor_create(n, ro.data(), ci.data(), va.data(), S.perm.data(), S.nsuper,
                                      S.sup_ptr.data(), 64, 1e-8);
    check(F != nullptr, "orc_factor_create", ordering);
    if (F) {
      std::vector<double> x(n), r(n), xb(n);
      orc_factor_solve(F, b.data(), x.data());
      const int st = orc_refine(F, n, ro.data(), ci.data(), va.data(), b.data(), x.data(), 1e-12, 10);
      check(st >= 0, "orc_refine", ordering);
      orc_spmv_upper(n, ro.data(), ci.data(), va.data(), x.data(), r.data());
      for (int i = 0; i < n; i++) r[i] = b[i] - r[i];
      resid = norm2(r) / norm2(b);
      const int it = orc_bicgstab(F, n, ro.data(), ci.data(), va.data(), b.data(), xb.data(), 1e-10, 30);
      check(it >= 1, "orc_bicgstab", ordering);
      long info[5];
      orc_factor_info(F, info);
      check(info[0] + info[1] == n, "inertia sums to n", ordering);
      orc_factor_free(F);
    }
    const int ns = S.nsuper;
    const int64_t nnzL = S.nnzL;
    std::fwrite(&ns, 4, 1, o);
    std::fwrite(S.perm.data(), 4, n, o);
    std::fwrite(S.sup_ptr.data(), 4, ns + 1, o);
    std::fwrite(&nnzL, 8, 1, o);
    std::fwrite(&resid, 8, 1, o);
  }
  std::fclose(o);
  return fails ? 1 : 0;
}

// malformed upper-CSR patterns must be rejected with an error, never read out
// of bounds; patterns the convention tolerates (entries below the diagonal,
// which are ignored; unsorted rows; a row with no diagonal entry) must be
// accepted and analysed in both orderings
int run_bad() {
  struct Case {
    const char* what;
    int n;
    std::vector<int> ro, ci;
  };
  const std::vector<Case> reject = {
      {"column out of range", 3, {0, 2, 3, 4}, {0, 3, 1, 2}},
      {"negative column", 3, {0, 2, 3, 4}, {0, -1, 1, 2}},
      {"row offsets decreasing", 3, {0, 2, 1, 4}, {0, 1, 1, 2}},
      {"row offsets past nnz", 3, {0, 2, 3, 9}, {0, 1, 1, 2}},
      {"first row offset not 0", 3, {1, 2, 3, 4}, {0, 1, 1, 2}},
      {"duplicate column", 3, {0, 3, 4, 5}, {0, 1, 1, 1, 2}},
  };
  const std::vector<Case> accept = {
      {"entry below the diagonal", 3, {0, 1, 3, 4}, {0, 0, 1, 2}},
      {"unsorted columns", 4, {0, 3, 5, 6, 7}, {2, 0, 1, 3, 1, 2, 3}},
      {"row without a diagonal entry", 4, {0, 2, 3, 4, 5}, {0, 3, 2, 2, 3}},
  };
  for (const Case& c : reject) {
    std::string err;
    if (kkt::validate_pattern(c.n, (int)c.ci.size(), c.ro.data(), c.ci.data(), err) == 0) {
      std::fprintf(stderr, "FAIL: accepted %s\n", c.what);
      fails++;
    }
  }
  for (const Case& c : accept) {
    std::string err;
    if (kkt::validate_pattern(c.n, (int)c.ci.size(), c.ro.data(), c.ci.data(), err) != 0) {
      std::fprintf(stderr, "FAIL: rejected %s (%s)\n", c.what, err.c_str());
      fails++;
      continue;
    }
    for (int ordering = 0; ordering < 2; ordering++) {
      kkt::SymOptions opt;
      opt.ordering = ordering;
      kkt::Symbolic S;
      if (kkt::symbolic_analyze(c.n, c.ro.data(), c.ci.data(), nullptr, opt, S, err) != 0 ||
          S.sup_ptr.empty() || S.sup_ptr.back() != c.n) {
        std::fprintf(stderr, "FAIL: analysis of %s (ordering %d)\n", c.what, ordering);
        fails++;
      }
    }
  }
  return fails ? 1 : 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc == 2 && std::strcmp(argv[1], "--bad") == 0) return run_bad();
  if (argc != 3) {
    std::fprintf(stderr, "usage: san_driver IN.bin OUT.bin | san_driver --bad\n");
    return 2;
  }
  return run_matrix(argv[1], argv[2]);
}
