// table_check.cpp -- test-only: every FILTER line of a table is a member of the family it
// names, with a modulus that does not divide the wheel (so it is not already spent on the
// wheel): an F1 cover (m, res) from covers_upto(mmax), or an F2 uniform rung certificate
// (L, res) from uniform_certs(63, 64). A table made only of such members can leave extra
// survivors but cannot certify a class wrongly.
//
// usage: table_check TABLE M MMAX     prints "ok FILTERS" or the first offending line.
#include "es_plan.h"

#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <set>
#include <utility>

using namespace es;

int main(int argc, char **argv) {
  if (argc != 4) return 2;
  const u64 M = strtoull(argv[2], 0, 10), mmax = strtoull(argv[3], 0, 10);
  std::set<std::pair<u64, u64>> f1, f2;
  for (const Cover &c : covers_upto(mmax)) f1.insert({c.m, c.res});
  for (const Ucert &u : uniform_certs(kUcertRmax, kUcertUmax)) f2.insert({u.L, u.res});
  FILE *f = fopen(argv[1], "r");
  if (!f) return 2;
  char line[256];
  u64 n = 0;
  while (fgets(line, sizeof line, f)) {
    u64 mod, res;
    char kind[8] = {0};
    if (sscanf(line, "FILTER %" SCNu64 " %" SCNu64 " %7s", &mod, &res, kind) != 3) continue;
    const bool ok = M % mod != 0 && ((!strcmp(kind, "F1") && f1.count({mod, res})) ||
                                     (!strcmp(kind, "F2") && f2.count({mod, res})));
    if (!ok) {
      printf("not certified: %s", line);
      return 1;
    }
    ++n;
  }
  printf("ok %" PRIu64 "\n", n);
  return 0;
}
