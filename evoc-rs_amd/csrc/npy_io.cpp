// .npy / .npz arrays: the on-disk format of the reference's fixtures, goldens and
// bench inputs (read there with ndarray-npy: tests/parity.rs:22-30,
// benches/evoc_bench.rs:4-15, src/graph_construction.rs:117-140, where
// NpzReader::by_name takes a member's name with or without ".npy").  Host-only
// code: little-endian, C-order arrays of the
// element types the kNN path and its fixtures use; .npz members stored without
// compression (numpy.savez), found through the zip central directory (zip64
// fields included, as numpy writes them for large members).
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "evoc_mi355x.h"

namespace {

struct ElemType {
  const char* descr;  // numpy descr without the byte-order character
  int32_t code;
  int32_t size;
};

const ElemType kTypes[] = {
    {"b1", EVOC_NPY_BOOL, 1}, {"i1", EVOC_NPY_I8, 1},  {"u1", EVOC_NPY_U8, 1},
    {"i4", EVOC_NPY_I32, 4},  {"u4", EVOC_NPY_U32, 4}, {"i8", EVOC_NPY_I64, 8},
    {"u8", EVOC_NPY_U64, 8},  {"f4", EVOC_NPY_F32, 4}, {"f8", EVOC_NPY_F64, 8},
};

const ElemType* type_of_code(int32_t code) {
  for (const auto& t : kTypes)
    if (t.code == code) return &t;
  return nullptr;
}

struct Err {
  char* buf;
  int64_t len;
  int fail(int code, const std::string& msg) const {
    if (buf && len > 0) {
      strncpy(buf, msg.c_str(), (size_t)len - 1);
      buf[len - 1] = '\0';
    }
    return code;
  }
};

uint64_t rd_le(const uint8_t* p, int bytes) {
  uint64_t v = 0;
  for (int i = bytes - 1; i >= 0; i--) v = (v << 8) | p[i];
  return v;
}

struct File {
  FILE* f = nullptr;
  ~File() {
    if (f) fclose(f);
  }
  bool read_at(uint64_t off, void* dst, size_t n) {
    if (fseeko(f, (off_t)off, SEEK_SET) != 0) return false;
    return fread(dst, 1, n, f) == n;
  }
  uint64_t size() {
    fseeko(f, 0, SEEK_END);
    return (uint64_t)ftello(f);
  }
};

// Python-literal header dict: {'descr': '<f4', 'fortran_order': False, 'shape': (3, 4), }
int parse_header(const std::string& h, evoc_npy_array* a, int32_t* elem_size, const Err& err) {
  auto value_after = [&](const char* key) -> size_t {
    const std::string k = std::string("'") + key + "'";
    size_t p = h.find(k);
    if (p == std::string::npos) return std::string::npos;
    p = h.find(':', p + k.size());
    if (p == std::string::npos) return std::string::npos;
    p++;
    while (p < h.size() && h[p] == ' ') p++;
    return p;
  };
  size_t p = value_after("descr");
  if (p == std::string::npos || p >= h.size() || (h[p] != '\'' && h[p] != '"'))
    return err.fail(EVOC_ERR_INVALID_ARGUMENT, "npy header: no 'descr' string");
  const size_t q = h.find(h[p], p + 1);
  if (q == std::string::npos) return err.fail(EVOC_ERR_INVALID_ARGUMENT, "npy header: unterminated 'descr'");
  const std::string descr = h.substr(p + 1, q - p - 1);
  if (descr.size() != 3 || (descr[0] != '<' && descr[0] != '|' && descr[0] != '='))
    return err.fail(EVOC_ERR_UNSUPPORTED, "npy dtype '" + descr + "': little-endian scalar types only");
  const ElemType* t = nullptr;
  for (const auto& c : kTypes)
    if (descr.compare(1, 2, c.descr) == 0) t = &c;
  if (!t) return err.fail(EVOC_ERR_UNSUPPORTED, "npy dtype '" + descr + "' is not supported");
  p = value_after("fortran_order");
  if (p == std::string::npos) return err.fail(EVOC_ERR_INVALID_ARGUMENT, "npy header: no 'fortran_order'");
  if (h.compare(p, 4, "True") == 0)
    return err.fail(EVOC_ERR_UNSUPPORTED, "npy array in Fortran order: C order only");
  if (h.compare(p, 5, "False") != 0) return err.fail(EVOC_ERR_INVALID_ARGUMENT, "npy header: bad 'fortran_order'");
  p = value_after("shape");
  if (p == std::string::npos || h[p] != '(') return err.fail(EVOC_ERR_INVALID_ARGUMENT, "npy header: no 'shape' tuple");
  const size_t e = h.find(')', p);
  if (e == std::string::npos) return err.fail(EVOC_ERR_INVALID_ARGUMENT, "npy header: unterminated 'shape'");
  int nd = 0;
  size_t i = p + 1;
  while (i < e) {
    while (i < e && (h[i] == ' ' || h[i] == ',')) i++;
    if (i >= e) break;
    if (h[i] < '0' || h[i] > '9') return err.fail(EVOC_ERR_INVALID_ARGUMENT, "npy header: bad 'shape'");
    int64_t v = 0;
    while (i < e && h[i] >= '0' && h[i] <= '9') {
      if (v > (INT64_MAX - 9) / 10) return err.fail(EVOC_ERR_INVALID_ARGUMENT, "npy header: dimension overflows");
      v = v * 10 + (h[i] - '0');
      i++;
    }
    if (nd >= EVOC_NPY_MAX_DIMS) return err.fail(EVOC_ERR_UNSUPPORTED, "npy array has more than 8 dimensions");
    a->shape[nd++] = v;
  }
  a->ndim = nd;
  a->dtype = t->code;
  *elem_size = t->size;
  return EVOC_OK;
}

// Parses the array whose .npy image starts at byte `base` of the file and spans
// `limit` bytes (the whole rest of the file for a plain .npy).
int read_npy_at(File& f, uint64_t base, uint64_t limit, evoc_npy_array* out, const Err& err) {
  uint8_t pre[12];
  if (limit < 10 || !f.read_at(base, pre, 10)) return err.fail(EVOC_ERR_INVALID_ARGUMENT, "npy: truncated preamble");
  if (memcmp(pre, "\x93NUMPY", 6) != 0) return err.fail(EVOC_ERR_INVALID_ARGUMENT, "npy: bad magic");
  const int major = pre[6];
  uint64_t hlen, hstart;
  if (major == 1) {
    hlen = rd_le(pre + 8, 2);
    hstart = 10;
  } else if (major == 2 || major == 3) {
    if (limit < 12 || !f.read_at(base, pre, 12)) return err.fail(EVOC_ERR_INVALID_ARGUMENT, "npy: truncated preamble");
    hlen = rd_le(pre + 8, 4);
    hstart = 12;
  } else {
    return err.fail(EVOC_ERR_UNSUPPORTED, "npy format version " + std::to_string(major) + " is not supported");
  }
  if (hstart + hlen > limit) return err.fail(EVOC_ERR_INVALID_ARGUMENT, "npy: truncated header");
  std::string h(hlen, '\0');
  if (!f.read_at(base + hstart, &h[0], hlen)) return err.fail(EVOC_ERR_INVALID_ARGUMENT, "npy: truncated header");
  evoc_npy_array a;
  memset(&a, 0, sizeof(a));
  int32_t esz = 0;
  int rc = parse_header(h, &a, &esz, err);
  if (rc != EVOC_OK) return rc;
  uint64_t count = 1;
  for (int d = 0; d < a.ndim; d++) {
    if (a.shape[d] < 0) return err.fail(EVOC_ERR_INVALID_ARGUMENT, "npy: negative dimension");
    if (a.shape[d] != 0 && count > UINT64_MAX / (uint64_t)a.shape[d])
      return err.fail(EVOC_ERR_INVALID_ARGUMENT, "npy: element count overflows");
    count *= (uint64_t)a.shape[d];
  }
  if (count > UINT64_MAX / (uint64_t)esz) return err.fail(EVOC_ERR_INVALID_ARGUMENT, "npy: byte count overflows");
  a.nbytes = count * (uint64_t)esz;
  if (a.nbytes > limit - (hstart + hlen))
    return err.fail(EVOC_ERR_INVALID_ARGUMENT, "npy: data shorter than its shape (" + std::to_string(a.nbytes) +
                                                   " bytes expected)");
  a.data = malloc(a.nbytes ? a.nbytes : 1);
  if (!a.data) return err.fail(EVOC_ERR_OUT_OF_BUDGET, "npy: cannot allocate " + std::to_string(a.nbytes) + " bytes");
  if (a.nbytes && !f.read_at(base + hstart + hlen, a.data, a.nbytes)) {
    free(a.data);
    return err.fail(EVOC_ERR_INVALID_ARGUMENT, "npy: read failed");
  }
  *out = a;
  return EVOC_OK;
}

int open_file(File& f, const char* path, const char* mode, const Err& err) {
  if (!path) return err.fail(EVOC_ERR_INVALID_ARGUMENT, "null path");
  f.f = fopen(path, mode);
  if (!f.f) return err.fail(EVOC_ERR_INVALID_ARGUMENT, std::string("cannot open ") + path);
  return EVOC_OK;
}

}  // namespace

extern "C" {

int evoc_npy_read(const char* path, evoc_npy_array* out, char* err, int64_t err_len) {
  const Err e{err, err_len};
  if (!out) return e.fail(EVOC_ERR_INVALID_ARGUMENT, "null output");
  memset(out, 0, sizeof(*out));
  File f;
  int rc = open_file(f, path, "rb", e);
  if (rc != EVOC_OK) return rc;
  return read_npy_at(f, 0, f.size(), out, e);
}

int evoc_npz_read(const char* path, const char* member, evoc_npy_array* out, char* err, int64_t err_len) {
  const Err e{err, err_len};
  if (!out || !member) return e.fail(EVOC_ERR_INVALID_ARGUMENT, "null member or output");
  memset(out, 0, sizeof(*out));
  File f;
  int rc = open_file(f, path, "rb", e);
  if (rc != EVOC_OK) return rc;
  const uint64_t fsize = f.size();
  // end of central directory: 22 bytes plus a comment of up to 65535
  const uint64_t tail = fsize < 22 + 65535 ? fsize : 22 + 65535;
  std::vector<uint8_t> t(tail);
  if (tail < 22 || !f.read_at(fsize - tail, t.data(), tail)) return e.fail(EVOC_ERR_INVALID_ARGUMENT, "npz: not a zip file");
  int64_t eocd = -1;
  for (int64_t i = (int64_t)tail - 22; i >= 0; i--)
    if (rd_le(&t[i], 4) == 0x06054b50u) {
      eocd = i;
      break;
    }
  if (eocd < 0) return e.fail(EVOC_ERR_INVALID_ARGUMENT, "npz: no end-of-central-directory record");
  uint64_t n_entries = rd_le(&t[eocd + 10], 2), cd_size = rd_le(&t[eocd + 12], 4), cd_off = rd_le(&t[eocd + 16], 4);
  if (n_entries == 0xffff || cd_size == 0xffffffffu || cd_off == 0xffffffffu) {
    // zip64: the locator sits right before the record and points at the zip64 record
    const uint64_t eocd_abs = fsize - tail + (uint64_t)eocd;
    uint8_t loc[20], rec[56];
    if (eocd_abs < 20 || !f.read_at(eocd_abs - 20, loc, 20) || rd_le(loc, 4) != 0x07064b50u)
      return e.fail(EVOC_ERR_INVALID_ARGUMENT, "npz: missing zip64 locator");
    if (!f.read_at(rd_le(loc + 8, 8), rec, 56) || rd_le(rec, 4) != 0x06064b50u)
      return e.fail(EVOC_ERR_INVALID_ARGUMENT, "npz: bad zip64 end record");
    n_entries = rd_le(rec + 32, 8);
    cd_size = rd_le(rec + 40, 8);
    cd_off = rd_le(rec + 48, 8);
  }
  if (cd_off > fsize || cd_size > fsize - cd_off) return e.fail(EVOC_ERR_INVALID_ARGUMENT, "npz: central directory out of range");
  std::vector<uint8_t> cd(cd_size);
  if (cd_size && !f.read_at(cd_off, cd.data(), cd_size)) return e.fail(EVOC_ERR_INVALID_ARGUMENT, "npz: truncated central directory");
  const std::string want = member, want_npy = std::string(member) + ".npy";
  uint64_t p = 0;
  for (uint64_t i = 0; i < n_entries; i++) {
    if (p + 46 > cd_size || rd_le(&cd[p], 4) != 0x02014b50u) return e.fail(EVOC_ERR_INVALID_ARGUMENT, "npz: bad central directory entry");
    const int method = (int)rd_le(&cd[p + 10], 2);
    uint64_t csize = rd_le(&cd[p + 20], 4), usize = rd_le(&cd[p + 24], 4);
    const uint64_t nlen = rd_le(&cd[p + 28], 2), xlen = rd_le(&cd[p + 30], 2), clen = rd_le(&cd[p + 32], 2);
    uint64_t loff = rd_le(&cd[p + 42], 4);
    if (p + 46 + nlen + xlen + clen > cd_size) return e.fail(EVOC_ERR_INVALID_ARGUMENT, "npz: bad central directory entry");
    const std::string name((const char*)&cd[p + 46], nlen);
    // zip64 extra field: the 8-byte values of the fields saturated above, in order
    for (uint64_t x = p + 46 + nlen; x + 4 <= p + 46 + nlen + xlen;) {
      const uint64_t id = rd_le(&cd[x], 2), sz = rd_le(&cd[x + 2], 2);
      if (x + 4 + sz > p + 46 + nlen + xlen) break;
      if (id == 0x0001) {
        uint64_t y = x + 4;
        if (usize == 0xffffffffu && y + 8 <= x + 4 + sz) usize = rd_le(&cd[y], 8), y += 8;
        if (csize == 0xffffffffu && y + 8 <= x + 4 + sz) csize = rd_le(&cd[y], 8), y += 8;
        if (loff == 0xffffffffu && y + 8 <= x + 4 + sz) loff = rd_le(&cd[y], 8), y += 8;
      }
      x += 4 + sz;
    }
    p += 46 + nlen + xlen + clen;
    if (name != want && name != want_npy) continue;
    if (method != 0)
      return e.fail(EVOC_ERR_UNSUPPORTED, "npz member '" + name + "' is compressed: stored members only (numpy.savez)");
    if (csize != usize) return e.fail(EVOC_ERR_INVALID_ARGUMENT, "npz: stored member with differing sizes");
    uint8_t lh[30];
    if (!f.read_at(loff, lh, 30) || rd_le(lh, 4) != 0x04034b50u) return e.fail(EVOC_ERR_INVALID_ARGUMENT, "npz: bad local header");
    const uint64_t data = loff + 30 + rd_le(lh + 26, 2) + rd_le(lh + 28, 2);
    if (data > fsize || usize > fsize - data) return e.fail(EVOC_ERR_INVALID_ARGUMENT, "npz: member out of range");
    return read_npy_at(f, data, usize, out, e);
  }
  return e.fail(EVOC_ERR_INVALID_ARGUMENT, std::string("npz: no member '") + member + "'");
}

int evoc_npy_write(const char* path, const evoc_npy_array* a, char* err, int64_t err_len) {
  const Err e{err, err_len};
  if (!a) return e.fail(EVOC_ERR_INVALID_ARGUMENT, "null array");
  const ElemType* t = type_of_code(a->dtype);
  if (!t) return e.fail(EVOC_ERR_UNSUPPORTED, "unknown element type " + std::to_string(a->dtype));
  if (a->ndim < 0 || a->ndim > EVOC_NPY_MAX_DIMS) return e.fail(EVOC_ERR_INVALID_ARGUMENT, "ndim out of range");
  uint64_t count = 1;
  std::string shape = "(";
  for (int d = 0; d < a->ndim; d++) {
    if (a->shape[d] < 0) return e.fail(EVOC_ERR_INVALID_ARGUMENT, "negative dimension");
    count *= (uint64_t)a->shape[d];
    shape += std::to_string(a->shape[d]) + (a->ndim == 1 ? ",)" : d + 1 < a->ndim ? ", " : ")");
  }
  if (a->ndim == 0) shape += ")";
  const uint64_t nbytes = count * (uint64_t)t->size;
  if (a->nbytes != nbytes) return e.fail(EVOC_ERR_INVALID_ARGUMENT, "nbytes does not match the shape and type");
  if (nbytes && !a->data) return e.fail(EVOC_ERR_INVALID_ARGUMENT, "null data");
  const char order = t->size == 1 ? '|' : '<';
  std::string h = std::string("{'descr': '") + order + t->descr + "', 'fortran_order': False, 'shape': " + shape + ", }";
  // version 1.0 while the header fits 16 bits; the preamble and header pad to 64 bytes
  const bool v2 = h.size() + 1 + 10 > 65535;
  const size_t pre = v2 ? 12 : 10;
  const size_t total = (pre + h.size() + 1 + 63) / 64 * 64;
  h.append(total - pre - h.size() - 1, ' ');
  h.push_back('\n');
  uint8_t head[12] = {0x93, 'N', 'U', 'M', 'P', 'Y', (uint8_t)(v2 ? 2 : 1), 0};
  const uint64_t hl = h.size();
  for (size_t i = 0; i < (v2 ? 4u : 2u); i++) head[8 + i] = (uint8_t)(hl >> (8 * i));
  File f;
  int rc = open_file(f, path, "wb", e);
  if (rc != EVOC_OK) return rc;
  if (fwrite(head, 1, pre, f.f) != pre || fwrite(h.data(), 1, h.size(), f.f) != h.size() ||
      (nbytes && fwrite(a->data, 1, nbytes, f.f) != nbytes))
    return e.fail(EVOC_ERR_INVALID_ARGUMENT, std::string("write failed: ") + path);
  if (fclose(f.f) != 0) {
    f.f = nullptr;
    return e.fail(EVOC_ERR_INVALID_ARGUMENT, std::string("write failed: ") + path);
  }
  f.f = nullptr;
  return EVOC_OK;
}

void evoc_npy_free(evoc_npy_array* a) {
  if (!a) return;
  free(a->data);
  a->data = nullptr;
  a->nbytes = 0;
}

}  // extern "C"
