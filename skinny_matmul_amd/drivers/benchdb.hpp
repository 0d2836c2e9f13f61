// BenchDB: results of the perf driver in an sqlite database, one row per
// measured configuration (the reference's benchdb.hpp role and its
// `benchmarks` table, whose columns perf.cu:286-307 sets; the column names
// are kept so existing analysis over that table reads these rows too).
//
// * Rows are keyed by the configuration columns; inserting a key that is
//   already present replaces that row (latest measurement wins).
// * Inserts are queued and written in one transaction every 64 rows and at
//   destruction.  Values are bound as statement parameters, never spliced
//   into SQL text.
// * A second table, `tuning`, keeps the library's autotuner table per device
//   (the text of skinny_tune_export) so later runs start tuned.
//
// The image ships the sqlite runtime (libsqlite3.so.0) without its header;
// the handful of entry points used here are declared below with their
// documented C signatures, and the binaries link -l:libsqlite3.so.0.
#pragma once

#include <cstdint>
#include <cstdio>
#include <string>
#include <utility>
#include <vector>

extern "C" {
typedef struct sqlite3 sqlite3;
typedef struct sqlite3_stmt sqlite3_stmt;
int sqlite3_open(const char* filename, sqlite3** db);
int sqlite3_close(sqlite3* db);
int sqlite3_busy_timeout(sqlite3* db, int ms);
int sqlite3_exec(sqlite3* db, const char* sql,
                 int (*callback)(void*, int, char**, char**), void* arg,
                 char** errmsg);
const char* sqlite3_errmsg(sqlite3* db);
int sqlite3_prepare_v2(sqlite3* db, const char* sql, int nbyte,
                       sqlite3_stmt** stmt, const char** tail);
int sqlite3_bind_text(sqlite3_stmt* stmt, int idx, const char* text, int n,
                      void (*destructor)(void*));
int sqlite3_bind_int64(sqlite3_stmt* stmt, int idx, long long value);
int sqlite3_bind_double(sqlite3_stmt* stmt, int idx, double value);
int sqlite3_step(sqlite3_stmt* stmt);
int sqlite3_finalize(sqlite3_stmt* stmt);
const unsigned char* sqlite3_column_text(sqlite3_stmt* stmt, int col);
long long sqlite3_column_int64(sqlite3_stmt* stmt, int col);
double sqlite3_column_double(sqlite3_stmt* stmt, int col);
}

namespace drv {

// sqlite result codes and the "copy the text now" destructor
constexpr int kSqliteOk = 0;
constexpr int kSqliteRow = 100;
constexpr int kSqliteDone = 101;
inline void (*sqlite_transient())(void*) {
  return reinterpret_cast<void (*)(void*)>(static_cast<intptr_t>(-1));
}

// One column value: text, integer or real.
struct DbValue {
  enum Kind { TEXT, INT, REAL } kind;
  std::string text;
  long long i = 0;
  double d = 0.0;
  DbValue(const std::string& s) : kind(TEXT), text(s) {}
  DbValue(const char* s) : kind(TEXT), text(s) {}
  DbValue(int v) : kind(INT), i(v) {}
  DbValue(long long v) : kind(INT), i(v) {}
  DbValue(long v) : kind(INT), i(v) {}
  DbValue(unsigned long v) : kind(INT), i((long long)v) {}
  DbValue(double v) : kind(REAL), d(v) {}
};
typedef std::vector<std::pair<std::string, DbValue>> DbRow;

class BenchDB {
 public:
  // key and value column names of the `benchmarks` table
  static const std::vector<std::string>& key_columns() {
    static const std::vector<std::string> k = {
        "multype", "device",  "types",  "M",         "N",       "name",
        "inplace", "zerobeta", "branch", "usr1_name", "usr1_val"};
    return k;
  }
  static const std::vector<std::string>& value_columns() {
    static const std::vector<std::string> v = {
        "blockCount", "K",        "time",      "flops", "bw",
        "l2bw",       "sharedbw", "occupancy", "drambw"};
    return v;
  }

  explicit BenchDB(const std::string& path) {
    if (sqlite3_open(path.c_str(), &db_) != kSqliteOk) {
      std::fprintf(stderr, "BenchDB: cannot open %s: %s\n", path.c_str(),
                   db_ ? sqlite3_errmsg(db_) : "?");
      close();
      return;
    }
    sqlite3_busy_timeout(db_, 10000);
    exec(
        "CREATE TABLE IF NOT EXISTS benchmarks ("
        "multype TEXT, device TEXT, types TEXT, M INTEGER, N INTEGER, "
        "name TEXT, inplace INTEGER, zerobeta INTEGER, branch TEXT, "
        "usr1_name TEXT, usr1_val TEXT, blockCount INTEGER, K INTEGER, "
        "time REAL, flops REAL, bw REAL, l2bw REAL, sharedbw REAL, "
        "occupancy REAL, drambw REAL)");
    exec(
        "CREATE TABLE IF NOT EXISTS tuning (device TEXT PRIMARY KEY, "
        "table_text TEXT)");
  }
  ~BenchDB() {
    flush();
    close();
  }
  BenchDB(const BenchDB&) = delete;
  BenchDB& operator=(const BenchDB&) = delete;

  bool ok() const { return db_ != nullptr; }

  // keys: every key column; values: any subset of the value columns
  void insert(const DbRow& keys, const DbRow& values) {
    queue_.push_back({keys, values});
    if (queue_.size() >= 64) flush();
  }

  void flush() {
    if (!db_ || queue_.empty()) {
      queue_.clear();
      return;
    }
    exec("BEGIN TRANSACTION");
    for (const auto& e : queue_) write(e.first, e.second);
    exec("COMMIT TRANSACTION");
    queue_.clear();
  }

  void save_tuning(const std::string& device, const std::string& text) {
    if (!db_) return;
    sqlite3_stmt* st = prepare(
        "INSERT OR REPLACE INTO tuning (device, table_text) VALUES (?1, ?2)");
    if (!st) return;
    sqlite3_bind_text(st, 1, device.c_str(), -1, sqlite_transient());
    sqlite3_bind_text(st, 2, text.c_str(), -1, sqlite_transient());
    step_done(st);
  }

  // A stored measurement: the configuration to replay it.
  struct Stored {
    int M = 0, N = 0, inplace = 0, zerobeta = 0;
    std::string name, branch, usr1_name, usr1_val;
    long long blockCount = 0, K = 0;
    double time = 0.0;
  };
  // Rows of one operation, device and type set, in a fixed order.
  std::vector<Stored> rows(const std::string& multype, const std::string& device,
                           const std::string& types) {
    std::vector<Stored> out;
    if (!db_) return out;
    sqlite3_stmt* st = prepare(
        "SELECT M, N, name, inplace, zerobeta, blockCount, K, time, branch, usr1_name, "
        "usr1_val FROM benchmarks WHERE multype = ?1 AND device = ?2 AND types = ?3 "
        "ORDER BY M, N, name, inplace, zerobeta, branch, usr1_name, usr1_val");
    if (!st) return out;
    sqlite3_bind_text(st, 1, multype.c_str(), -1, sqlite_transient());
    sqlite3_bind_text(st, 2, device.c_str(), -1, sqlite_transient());
    sqlite3_bind_text(st, 3, types.c_str(), -1, sqlite_transient());
    while (sqlite3_step(st) == kSqliteRow) {
      Stored r;
      auto text = [&](int c) {
        const unsigned char* t = sqlite3_column_text(st, c);
        return t ? std::string(reinterpret_cast<const char*>(t)) : std::string();
      };
      r.M = (int)sqlite3_column_int64(st, 0);
      r.N = (int)sqlite3_column_int64(st, 1);
      r.name = text(2);
      r.inplace = (int)sqlite3_column_int64(st, 3);
      r.zerobeta = (int)sqlite3_column_int64(st, 4);
      r.blockCount = sqlite3_column_int64(st, 5);
      r.K = sqlite3_column_int64(st, 6);
      r.time = sqlite3_column_double(st, 7);
      r.branch = text(8);
      r.usr1_name = text(9);
      r.usr1_val = text(10);
      out.push_back(r);
    }
    sqlite3_finalize(st);
    return out;
  }

  std::string load_tuning(const std::string& device) {
    std::string out;
    if (!db_) return out;
    sqlite3_stmt* st = prepare("SELECT table_text FROM tuning WHERE device = ?1");
    if (!st) return out;
    sqlite3_bind_text(st, 1, device.c_str(), -1, sqlite_transient());
    if (sqlite3_step(st) == kSqliteRow) {
      const unsigned char* t = sqlite3_column_text(st, 0);
      if (t) out = reinterpret_cast<const char*>(t);
    }
    sqlite3_finalize(st);
    return out;
  }

 private:
  void close() {
    if (db_) sqlite3_close(db_);
    db_ = nullptr;
  }

  void exec(const char* sql) {
    if (!db_) return;
    char* err = nullptr;
    if (sqlite3_exec(db_, sql, nullptr, nullptr, &err) != kSqliteOk)
      std::fprintf(stderr, "BenchDB: %s: %s\n", sql, sqlite3_errmsg(db_));
  }

  sqlite3_stmt* prepare(const std::string& sql) {
    sqlite3_stmt* st = nullptr;
    if (sqlite3_prepare_v2(db_, sql.c_str(), -1, &st, nullptr) != kSqliteOk) {
      std::fprintf(stderr, "BenchDB: prepare %s: %s\n", sql.c_str(),
                   sqlite3_errmsg(db_));
      return nullptr;
    }
    return st;
  }

  void step_done(sqlite3_stmt* st) {
    if (sqlite3_step(st) != kSqliteDone)
      std::fprintf(stderr, "BenchDB: step: %s\n", sqlite3_errmsg(db_));
    sqlite3_finalize(st);
  }

  static void bind(sqlite3_stmt* st, int idx, const DbValue& v) {
    switch (v.kind) {
      case DbValue::TEXT:
        sqlite3_bind_text(st, idx, v.text.c_str(), -1, sqlite_transient());
        break;
      case DbValue::INT:
        sqlite3_bind_int64(st, idx, v.i);
        break;
      case DbValue::REAL:
        sqlite3_bind_double(st, idx, v.d);
        break;
    }
  }

  // delete the row with these keys, then insert keys + values
  void write(const DbRow& keys, const DbRow& values) {
    std::string del = "DELETE FROM benchmarks WHERE ";
    for (size_t i = 0; i < keys.size(); i++)
      del += (i ? " AND " : "") + keys[i].first + " IS ?" + std::to_string(i + 1);
    if (sqlite3_stmt* st = prepare(del)) {
      for (size_t i = 0; i < keys.size(); i++) bind(st, (int)i + 1, keys[i].second);
      step_done(st);
    }
    std::string cols, params;
    int n = 0;
    for (const auto* part : {&keys, &values})
      for (const auto& kv : *part) {
        cols += (n ? ", " : "") + kv.first;
        params += (n ? ", ?" : "?") + std::to_string(n + 1);
        n++;
      }
    std::string ins = "INSERT INTO benchmarks (" + cols + ") VALUES (" + params + ")";
    if (sqlite3_stmt* st = prepare(ins)) {
      int i = 1;
      for (const auto* part : {&keys, &values})
        for (const auto& kv : *part) bind(st, i++, kv.second);
      step_done(st);
    }
  }

  sqlite3* db_ = nullptr;
  std::vector<std::pair<DbRow, DbRow>> queue_;
};

}  // namespace drv
