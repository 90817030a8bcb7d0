// sdg_raw.cpp -- zstd-compressed raw state files in the reference's byte layout.
//
// File = 8-byte size_t header + one zstd frame (compression level 1) of the payload
// (RawBinary.cpp:42-72). The header holds ZSTD_compressBound(payload size), exactly as the
// reference writes it; its reader sizes the decompression buffer with it and uses the bytes it
// needs, so this reader reports the frame's true decompressed size instead.
//
// The library is linked against the system libzstd.so.1; the five functions used are declared
// here from zstd's stable API instead of through its header.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/sdg.h"
#include "sdg_internal.hpp"

extern "C" {
size_t ZSTD_compressBound(size_t srcSize);
size_t ZSTD_compress(void* dst, size_t dstCapacity, const void* src, size_t srcSize, int compressionLevel);
size_t ZSTD_decompress(void* dst, size_t dstCapacity, const void* src, size_t compressedSize);
unsigned ZSTD_isError(size_t code);
const char* ZSTD_getErrorName(size_t code);
}

namespace {

int rawFail(int code, const std::string& msg) {
  sdg::setGlobalError(msg);
  return code;
}

}  // namespace

namespace sdg {

int writeRawFile(const char* path, const void* data, size_t bytes) {
  if (!path || (!data && bytes)) return rawFail(SDG_E_INVALID, "raw file: null argument");
  const size_t bound = ZSTD_compressBound(bytes);
  std::vector<char> frame(bound);
  const size_t n = ZSTD_compress(frame.data(), bound, data, bytes, 1);
  if (ZSTD_isError(n)) return rawFail(SDG_E_INVALID, std::string("raw file: zstd: ") + ZSTD_getErrorName(n));
  FILE* f = std::fopen(path, "wb");
  if (!f) return rawFail(SDG_E_INVALID, std::string("raw file: cannot open ") + path + " for writing");
  const uint64_t header = bound;  // size_t on the reference's LP64 targets
  const bool ok = std::fwrite(&header, sizeof(header), 1, f) == 1 && std::fwrite(frame.data(), 1, n, f) == n;
  if (std::fclose(f) != 0 || !ok) return rawFail(SDG_E_INVALID, std::string("raw file: write failed: ") + path);
  return SDG_OK;
}

int readRawFile(const char* path, std::vector<char>& out) {
  if (!path) return rawFail(SDG_E_INVALID, "raw file: null path");
  FILE* f = std::fopen(path, "rb");
  if (!f) return rawFail(SDG_E_INVALID, std::string("raw file: cannot open ") + path);
  uint64_t header = 0;
  std::vector<char> frame;
  bool ok = std::fread(&header, sizeof(header), 1, f) == 1;
  if (ok) {
    char buf[1 << 16];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof(buf), f)) > 0) frame.insert(frame.end(), buf, buf + got);
  }
  std::fclose(f);
  if (!ok || header > (uint64_t(1) << 40)) return rawFail(SDG_E_INVALID, std::string("raw file: bad header: ") + path);
  out.assign(static_cast<size_t>(header), 0);
  const size_t n = ZSTD_decompress(out.data(), out.size(), frame.data(), frame.size());
  if (ZSTD_isError(n)) return rawFail(SDG_E_INVALID, std::string("raw file: zstd: ") + ZSTD_getErrorName(n));
  out.resize(n);
  return SDG_OK;
}

}  // namespace sdg

extern "C" {

int sdg_raw_file_write(const char* path, const void* payload, int64_t bytes) {
  if (bytes < 0) return rawFail(SDG_E_INVALID, "sdg_raw_file_write: negative size");
  return sdg::writeRawFile(path, payload, static_cast<size_t>(bytes));
}

int sdg_raw_file_read(const char* path, void* payload, int64_t capacity, int64_t* bytes) {
  std::vector<char> buf;
  const int rc = sdg::readRawFile(path, buf);
  if (rc) return rc;
  if (bytes) *bytes = static_cast<int64_t>(buf.size());
  if (!payload) return SDG_OK;  // size query
  if (capacity < static_cast<int64_t>(buf.size()))
    return rawFail(SDG_E_INVALID, "sdg_raw_file_read: buffer smaller than the payload");
  std::copy(buf.begin(), buf.end(), static_cast<char*>(payload));
  return SDG_OK;
}

}  // extern "C"
