// skinny_benchdb_check <db>: host-only exercise of BenchDB (no GPU).
// Writes three benchmark rows, replaces one of them by key, stores and reads
// back an autotuner table; the CPU test suite then reads the file with
// Python's sqlite3 and checks the contents.
#include <cstdio>

#include "benchdb.hpp"

int main(int argc, char** argv) {
  if (argc != 2) {
    std::fprintf(stderr, "usage: skinny_benchdb_check <db>\n");
    return 2;
  }
  drv::BenchDB db(argv[1]);
  if (!db.ok()) return 1;
  auto keys = [](int M, const char* name) {
    return drv::DbRow{{"multype", "TSMTTSM"}, {"device", "check"}, {"types", "DR"},
                      {"M", M}, {"N", M}, {"name", name}, {"inplace", 0},
                      {"zerobeta", 1}, {"branch", "main"}, {"usr1_name", ""},
                      {"usr1_val", ""}};
  };
  for (int M = 1; M <= 3; M++)
    db.insert(keys(M, "SKINNY"), {{"blockCount", 512}, {"K", 1LL << 26},
                                  {"time", 0.001 * M}, {"bw", 100.0 * M}});
  db.insert(keys(2, "SKINNY"), {{"blockCount", 1024}, {"K", 1LL << 26},
                                {"time", 0.5}, {"bw", 4321.0}});
  db.flush();
  db.save_tuning("check", "0 0 8 8 26 1 512\n");
  const std::string t = db.load_tuning("check");
  std::printf("tuning: %s", t.c_str());
  return t == "0 0 8 8 26 1 512\n" ? 0 : 1;
}
