// fippi_cache.cpp -- process-wide FIPPI record cache and its file format
// (FIPPICache.cc:189-414); see fippi_cache.h for the layout.
#include "fippi_cache.h"

#include <sys/stat.h>

#include <cstdio>
#include <mutex>
#include <unordered_map>
#include <vector>

namespace scuffx {

namespace {

const char kSignature[] = "FIPPICACHE";            // 11 bytes with the terminator
const size_t kSigSize = sizeof(kSignature);
const size_t kRecSize = 352;                        // 60 + 4 + 288
const size_t kDataOff = 64;

struct Store {
  std::mutex mu;
  std::unordered_map<Key15, size_t, Key15Hash> index;   // key -> slot
  std::vector<Key15> keys;                               // insertion order
  std::vector<double> recs;                              // 36 per slot
  std::string preload_file;                              // last preload and its size
  size_t preloaded = 0;
};

Store &store()
{
  static Store s;
  return s;
}

}  // namespace

Key15 fippi_key(const double *const OVa[3], const double *const OVb[3])
{
  Key15 K;
  const double *src[5] = {OVa[1], OVa[2], OVb[0], OVb[1], OVb[2]};
  for (int s = 0; s < 5; s++)
    for (int m = 0; m < 3; m++) K.k[3 * s + m] = ((float)src[s][m]) - ((float)OVa[0][m]);
  return K;
}

bool fippi_cache_find(const Key15 &K, double *rec)
{
  Store &S = store();
  std::lock_guard<std::mutex> g(S.mu);
  auto it = S.index.find(K);
  if (it == S.index.end()) return false;
  std::memcpy(rec, S.recs.data() + it->second * SCUFFX_FIPPI_RECLEN, sizeof(double) * SCUFFX_FIPPI_RECLEN);
  return true;
}

static void insert_locked(Store &S, const Key15 &K, const double *rec)
{
  if (S.index.count(K)) return;
  S.index.emplace(K, S.keys.size());
  S.keys.push_back(K);
  S.recs.insert(S.recs.end(), rec, rec + SCUFFX_FIPPI_RECLEN);
}

void fippi_cache_insert(const Key15 &K, const double *rec)
{
  Store &S = store();
  std::lock_guard<std::mutex> g(S.mu);
  insert_locked(S, K, rec);
}

size_t fippi_cache_size()
{
  Store &S = store();
  std::lock_guard<std::mutex> g(S.mu);
  return S.keys.size();
}

void fippi_cache_clear()
{
  Store &S = store();
  std::lock_guard<std::mutex> g(S.mu);
  S.index.clear();
  S.keys.clear();
  S.recs.clear();
  S.preload_file.clear();
  S.preloaded = 0;
}

// FIPPICache::PreLoad (FIPPICache.cc:304-400): signature, size a whole
// number of records, then every record inserted (existing keys kept)
long fippi_cache_preload(const char *path, std::string &err)
{
  FILE *f = std::fopen(path, "rb");
  if (!f) { err = std::string("could not open ") + path; return -1; }
  struct stat st;
  if (fstat(fileno(f), &st)) { std::fclose(f); err = "invalid cache file"; return -1; }
  long long size = (long long)st.st_size;
  char sig[kSigSize];
  if (size < (long long)kSigSize || std::fread(sig, kSigSize, 1, f) != 1 || std::memcmp(sig, kSignature, kSigSize)) {
    std::fclose(f);
    err = "invalid cache file";
    return -1;
  }
  size -= (long long)kSigSize;
  if (size % (long long)kRecSize) { std::fclose(f); err = "cache file has incorrect size"; return -1; }
  const size_t n = (size_t)(size / (long long)kRecSize);
  std::vector<unsigned char> buf(kRecSize * n);
  if (n && std::fread(buf.data(), kRecSize, n, f) != n) { std::fclose(f); err = "short read"; return -1; }
  std::fclose(f);
  Store &S = store();
  std::lock_guard<std::mutex> g(S.mu);
  for (size_t r = 0; r < n; r++) {
    Key15 K;
    double rec[SCUFFX_FIPPI_RECLEN];
    std::memcpy(K.k, buf.data() + r * kRecSize, sizeof(K.k));
    std::memcpy(rec, buf.data() + r * kRecSize + kDataOff, sizeof(rec));
    insert_locked(S, K, rec);
  }
  S.preload_file = path;
  S.preloaded = S.keys.size();
  return (long)n;
}

// FIPPICache::Store (FIPPICache.cc:230-302): skipped when the cache was
// preloaded from the same file and has not grown since
long fippi_cache_store(const char *path, std::string &err)
{
  Store &S = store();
  std::lock_guard<std::mutex> g(S.mu);
  if (!S.preload_file.empty() && S.preload_file == path && S.preloaded == S.keys.size()) return 0;
  FILE *f = std::fopen(path, "wb");
  if (!f) { err = std::string("could not open ") + path; return -1; }
  bool ok = std::fwrite(kSignature, kSigSize, 1, f) == 1;
  unsigned char rec[kRecSize];
  long written = 0;
  for (size_t r = 0; ok && r < S.keys.size(); r++) {
    std::memset(rec, 0, sizeof(rec));
    std::memcpy(rec, S.keys[r].k, sizeof(S.keys[r].k));
    std::memcpy(rec + kDataOff, S.recs.data() + r * SCUFFX_FIPPI_RECLEN, sizeof(double) * SCUFFX_FIPPI_RECLEN);
    ok = std::fwrite(rec, kRecSize, 1, f) == 1;
    if (ok) written++;
  }
  if (std::fclose(f) != 0) ok = false;
  if (!ok) { err = std::string("write failed: ") + path; return -1; }
  return written;
}

}  // namespace scuffx
