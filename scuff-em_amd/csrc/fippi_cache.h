// fippi_cache.h -- process-wide cache of frequency-independent panel-pair
// integrals (FIPPI records) and its file format, the counterpart of the
// reference's GlobalFIPPICache with PreloadCache / StoreCache
// (libs/libscuff/FIPPICache.cc:46-414, libscuff.h:662-663).
//
// Key: 15 float32 values from the canonically ordered vertices,
// (float)V - (float)OVa[0] for V = OVa[1], OVa[2], OVb[0], OVb[1], OVb[2]
// (FIPPICache.cc:148-167).  Record: the 36 doubles of QIFIPPIData
// (libscuffInternals.h:282-292).  File: the 11 bytes "FIPPICACHE\0",
// then records of 352 bytes each: key (60 bytes), 4 bytes of padding,
// record (288 bytes) -- the in-memory layout of the reference's
// FIPPICF_Record on x86-64 (FIPPICache.cc:208-228), little-endian.
#pragma once

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <string>

namespace scuffx {

#define SCUFFX_FIPPI_KEYLEN 15
#define SCUFFX_FIPPI_RECLEN 36

struct Key15 {
  float k[SCUFFX_FIPPI_KEYLEN];
  bool operator==(const Key15 &o) const { return std::memcmp(k, o.k, sizeof(k)) == 0; }
};
struct Key15Hash {
  size_t operator()(const Key15 &K) const
  {
    const unsigned char *p = (const unsigned char *)K.k;
    uint64_t h = 1469598103934665603ULL;
    for (size_t i = 0; i < sizeof(K.k); i++) { h ^= p[i]; h *= 1099511628211ULL; }
    return (size_t)h;
  }
};

// key of a canonically ordered panel pair (OVa, OVb: 3 vertices each)
Key15 fippi_key(const double *const OVa[3], const double *const OVb[3]);

// the process-wide cache; insert keeps an existing record for the key
bool fippi_cache_find(const Key15 &K, double *rec);
void fippi_cache_insert(const Key15 &K, const double *rec);
size_t fippi_cache_size();
void fippi_cache_clear();
// file I/O; return the number of records read / written, or -1 with err set
long fippi_cache_preload(const char *path, std::string &err);
long fippi_cache_store(const char *path, std::string &err);

}  // namespace scuffx
