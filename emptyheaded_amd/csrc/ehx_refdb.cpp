// Reader for relations the reference engine saved to disk, so a database
// directory its loader wrote opens here without rebuilding it.
//
// Layout (storage_engine/src/Trie.cpp:15-35 save, :226-377 build;
// src/utils/ParMemoryBuffer.cpp:64-99; src/utils/MemoryBuffer.cpp:104-170):
//   <relation>/<folder>/trieinfo.bin  (folder "ram" for ParMemoryBuffer, "mmap"
//       for ParMMapBuffer): bool annotated, A annotation, size_t num_rows,
//       size_t num_columns, size_t num_buffers, size_t head size, then one
//       size_t per thread buffer;
//   <relation>/<folder>/data_head.bin, data_<i>.bin: the bump buffers; under
//       ram/ each is "<size> <offset> " in decimal text followed by <size> raw
//       bytes, under mmap/ the file is the buffer itself
//       (src/utils/ParMMapBuffer.cpp:1-40, src/utils/MMapBuffer.cpp:25-46) and
//       trieinfo.bin gives its used size.
// The head TrieBlock sits at offset 0 of the head buffer (Trie.cpp:120-125).
// A TrieBlock is one byte, followed by the 24-byte Set header (card u32,
// range u32 = largest element, number_of_bytes u64, type u8), the set's bytes
// (uinteger: card u32 values; range_bitset: u64 first word index, W u64 words,
// W u32 ranks, layouts/range_bitset.hpp:123-157), then, below the last level,
// NextLevel[nextSize] = {int buffer index (-1: none), size_t offset}, or on
// the last level of an annotated relation A[nextSize].  nextSize and the slot
// of element (i, v) follow TrieBlock.hpp:26-40: by position i when the set is
// sparse by common::is_sparse(card, range), else by value v.  Buffer index
// num_buffers is the head buffer (ParMemoryBuffer.cpp:83).
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "ehx_internal.h"

struct ehx_refdb {
  uint32_t ncols = 0;
  int annotated = 0;
  uint64_t rows = 0;
  std::vector<std::vector<uint32_t>> cols;
  std::vector<uint64_t> annos;
};

namespace ehx {
namespace {

struct Image {
  std::vector<std::vector<uint8_t>> bufs;  // thread buffers, then the head buffer
  size_t anno_bytes = 8;
  uint32_t ncols = 0;
  bool annotated = false;
};

bool read_file(const std::string &path, std::vector<uint8_t> &out) {
  FILE *f = fopen(path.c_str(), "rb");
  if (!f) return false;
  out.clear();
  uint8_t tmp[1 << 16];
  size_t k;
  while ((k = fread(tmp, 1, sizeof(tmp), f)) > 0) out.insert(out.end(), tmp, tmp + k);
  fclose(f);
  return true;
}

// "<size> <offset> " then size bytes (MemoryBuffer::save / load).
int parse_buffer(const std::string &path, std::vector<uint8_t> &buf) {
  std::vector<uint8_t> raw;
  if (!read_file(path, raw)) return fail(EHX_ERR_INVALID, "cannot open %s", path.c_str());
  size_t p = 0;
  auto number = [&](uint64_t &v) {
    while (p < raw.size() && (raw[p] == ' ' || raw[p] == '\n' || raw[p] == '\t' || raw[p] == '\r')) p++;
    if (p >= raw.size() || raw[p] < '0' || raw[p] > '9') return false;
    v = 0;
    while (p < raw.size() && raw[p] >= '0' && raw[p] <= '9') v = v * 10 + (raw[p++] - '0');
    return true;
  };
  uint64_t size = 0, offset = 0;
  if (!number(size) || !number(offset) || p >= raw.size())
    return fail(EHX_ERR_INVALID, "%s: no buffer header", path.c_str());
  p++;  // the single separator the loader skips with ifile.get()
  if (raw.size() - p < size) return fail(EHX_ERR_INVALID, "%s: %llu bytes announced, %llu present", path.c_str(),
                                         (unsigned long long)size, (unsigned long long)(raw.size() - p));
  buf.assign(raw.begin() + p, raw.begin() + p + size);
  return EHX_OK;
}

struct SetView {
  uint32_t card, range;
  uint64_t nbytes;
  uint8_t type;
  size_t data;  // byte offset of the set's data in its buffer
};

int read_set(const std::vector<uint8_t> &b, size_t block, SetView &s) {
  if (block > b.size() || b.size() - block < 1 + 24)
    return fail(EHX_ERR_INVALID, "TrieBlock at %zu past its buffer", block);
  const uint8_t *h = b.data() + block + 1;
  memcpy(&s.card, h, 4);
  memcpy(&s.range, h + 4, 4);
  memcpy(&s.nbytes, h + 8, 8);
  s.type = h[16];
  s.data = block + 1 + 24;
  if (s.nbytes > b.size() - s.data) return fail(EHX_ERR_INVALID, "set at %zu: %llu bytes past its buffer", block,
                                                (unsigned long long)s.nbytes);
  return EHX_OK;
}

int decode_set(const std::vector<uint8_t> &b, const SetView &s, std::vector<uint32_t> &v) {
  v.clear();
  if (s.card == 0) return EHX_OK;
  const uint8_t *d = b.data() + s.data;
  if (s.type == EHX_UINTEGER) {
    if (s.nbytes < 4ull * s.card) return fail(EHX_ERR_INVALID, "uinteger set shorter than its cardinality");
    v.resize(s.card);
    memcpy(v.data(), d, 4ull * s.card);
  } else if (s.type == EHX_RANGE_BITSET) {
    if (s.nbytes < 8) return fail(EHX_ERR_INVALID, "range_bitset set without its first word");
    uint64_t start;
    memcpy(&start, d, 8);
    const uint64_t W = (s.nbytes - 8) / 12;
    v.reserve(s.card);
    for (uint64_t w = 0; w < W; w++) {
      uint64_t word;
      memcpy(&word, d + 8 + 8 * w, 8);
      while (word) {
        const int bit = __builtin_ctzll(word);
        word &= word - 1;
        v.push_back((uint32_t)((start + w) * 64 + bit));
      }
    }
    if (v.size() != s.card) return fail(EHX_ERR_INVALID, "range_bitset holds %zu values, header says %u", v.size(),
                                        s.card);
  } else {
    return fail(EHX_ERR_UNSUPPORTED, "set layout %u", (unsigned)s.type);
  }
  return EHX_OK;
}

inline bool sparse(uint64_t card, uint64_t range) { return card <= 2 || range > 128 * card; }

// Walks the block at (buffer, offset) on `level`, appending rows under prefix.
int walk(const Image &im, size_t buf, size_t off, uint32_t level, std::vector<uint32_t> &prefix, ehx_refdb *out) {
  if (buf >= im.bufs.size()) return fail(EHX_ERR_INVALID, "NextLevel names buffer %zu of %zu", buf, im.bufs.size());
  const std::vector<uint8_t> &b = im.bufs[buf];
  SetView s;
  int rc;
  if ((rc = read_set(b, off, s))) return rc;
  std::vector<uint32_t> vals;
  if ((rc = decode_set(b, s, vals))) return rc;
  const bool sp = sparse(s.card, s.range);
  const uint64_t next_size = sp ? s.card : (uint64_t)s.range + 1;
  const size_t tail = s.data + s.nbytes;
  // a dense set's NextLevel and annotation slots are indexed by value: every
  // value must lie in [0, range]
  if (!sp)
    for (size_t i = 0; i < vals.size(); i++)
      if (vals[i] > s.range)
        return fail(EHX_ERR_INVALID, "set at %zu: value %u above its range %u", off, vals[i], s.range);
  if (level + 1 == im.ncols) {
    if (im.annotated && next_size * im.anno_bytes > b.size() - tail)
      return fail(EHX_ERR_INVALID, "annotations past their buffer");
    for (size_t i = 0; i < vals.size(); i++) {
      for (uint32_t c = 0; c < level; c++) out->cols[c].push_back(prefix[c]);
      out->cols[level].push_back(vals[i]);
      if (im.annotated) {
        uint64_t a = 0;
        memcpy(&a, b.data() + tail + (sp ? i : vals[i]) * im.anno_bytes, im.anno_bytes);
        out->annos.push_back(a);
      }
    }
    return EHX_OK;
  }
  if (next_size * 16 > b.size() - tail) return fail(EHX_ERR_INVALID, "NextLevel array past its buffer");
  for (size_t i = 0; i < vals.size(); i++) {
    const uint8_t *e = b.data() + tail + 16 * (sp ? i : vals[i]);
    int32_t idx;
    uint64_t o;
    memcpy(&idx, e, 4);
    memcpy(&o, e + 8, 8);
    if (idx < 0) continue;  // no child set: no complete row through this value
    prefix[level] = vals[i];
    if ((rc = walk(im, (size_t)idx, (size_t)o, level + 1, prefix, out))) return rc;
  }
  return EHX_OK;
}

}  // namespace

int refdb_open(const char *dir, int anno_bytes, ehx_refdb **out) {
  if (!dir || !out || anno_bytes < 1 || anno_bytes > 8) return fail(EHX_ERR_INVALID, "bad argument");
  std::string base;
  std::vector<uint8_t> info;
  bool mmapped = false;
  for (const char *folder : {"/ram/", "/mmap/"}) {
    base = std::string(dir) + folder;
    if (read_file(base + "trieinfo.bin", info)) {
      mmapped = folder[1] == 'm';
      break;
    }
    base.clear();
  }
  if (base.empty()) return fail(EHX_ERR_INVALID, "%s: no ram/ or mmap/ trieinfo.bin", dir);
  Image im;
  im.anno_bytes = (size_t)anno_bytes;
  const size_t need = 1 + im.anno_bytes + 8 * 4;
  if (info.size() < need) return fail(EHX_ERR_INVALID, "%s: trieinfo.bin too short", dir);
  uint64_t num_rows, num_columns, num_buffers;
  im.annotated = info[0] != 0;
  memcpy(&num_rows, &info[1 + im.anno_bytes], 8);
  memcpy(&num_columns, &info[1 + im.anno_bytes + 8], 8);
  memcpy(&num_buffers, &info[1 + im.anno_bytes + 16], 8);
  if (num_columns > EHX_MAX_ATTRS || num_buffers > 4096)
    return fail(EHX_ERR_UNSUPPORTED, "%s: %llu columns, %llu buffers", dir, (unsigned long long)num_columns,
                (unsigned long long)num_buffers);
  im.ncols = (uint32_t)num_columns;
  im.bufs.resize(num_buffers + 1);
  int rc;
  if (mmapped) {  // raw files; used sizes: head, then each thread buffer
    const size_t sz0 = 1 + im.anno_bytes + 24;
    if (info.size() < sz0 + 8 * (num_buffers + 1)) return fail(EHX_ERR_INVALID, "%s: trieinfo.bin too short", dir);
    for (uint64_t i = 0; i <= num_buffers; i++) {
      uint64_t used;
      memcpy(&used, &info[sz0 + 8 * i], 8);
      const std::string f = i == 0 ? base + "data_head.bin" : base + "data_" + std::to_string(i - 1) + ".bin";
      std::vector<uint8_t> &dst = i == 0 ? im.bufs[num_buffers] : im.bufs[i - 1];
      if (!read_file(f, dst)) return fail(EHX_ERR_INVALID, "cannot open %s", f.c_str());
      if (dst.size() < used) return fail(EHX_ERR_INVALID, "%s: %llu bytes used, %zu present", f.c_str(),
                                         (unsigned long long)used, dst.size());
      dst.resize(used);
    }
  } else {
    for (uint64_t i = 0; i < num_buffers; i++)
      if ((rc = parse_buffer(base + "data_" + std::to_string(i) + ".bin", im.bufs[i]))) return rc;
    if ((rc = parse_buffer(base + "data_head.bin", im.bufs[num_buffers]))) return rc;
  }
  ehx_refdb *r = new ehx_refdb();
  r->ncols = im.ncols;
  r->annotated = im.annotated;
  r->cols.resize(im.ncols);
  if (im.ncols > 0) {
    std::vector<uint32_t> prefix(im.ncols);
    if ((rc = walk(im, num_buffers, 0, 0, prefix, r))) {
      delete r;
      return rc;
    }
    r->rows = r->cols[0].size();
  }
  *out = r;
  return EHX_OK;
}

// <dir>/encoded.bin (intermediate/EncodedColumnStore.hpp:88-178): size_t
// num_rows, num_attributes, num_annotations; per attribute a uint32 max set
// size and num_rows uint32 keys; per annotation a size_t byte width and
// num_rows values of that width.  Rows keep the file's order.
int encoded_open(const char *dir, ehx_refdb **out) {
  if (!dir || !out) return fail(EHX_ERR_INVALID, "bad argument");
  const std::string f = std::string(dir) + "/encoded.bin";
  std::vector<uint8_t> raw;
  if (!read_file(f, raw)) return fail(EHX_ERR_INVALID, "cannot open %s", f.c_str());
  uint64_t hdr[3];
  if (raw.size() < sizeof(hdr)) return fail(EHX_ERR_INVALID, "%s: too short", f.c_str());
  memcpy(hdr, raw.data(), sizeof(hdr));
  const uint64_t rows = hdr[0], na = hdr[1], nann = hdr[2];
  if (na > EHX_MAX_ATTRS || nann > 64) return fail(EHX_ERR_UNSUPPORTED, "%s: %llu attributes, %llu annotations",
                                                   f.c_str(), (unsigned long long)na, (unsigned long long)nann);
  size_t p = sizeof(hdr);
  ehx_refdb *r = new ehx_refdb();
  r->ncols = (uint32_t)na;
  r->rows = rows;
  r->cols.resize(na);
  for (uint64_t a = 0; a < na; a++) {
    if (raw.size() - p < 4 || (raw.size() - p - 4) / 4 < rows) {
      delete r;
      return fail(EHX_ERR_INVALID, "%s: column %llu truncated", f.c_str(), (unsigned long long)a);
    }
    p += 4;  // max set size
    r->cols[a].resize(rows);
    if (rows) memcpy(r->cols[a].data(), raw.data() + p, 4 * rows);
    p += 4 * rows;
  }
  for (uint64_t k = 0; k < nann; k++) {
    uint64_t width = 0;
    if (raw.size() - p < 8) {
      delete r;
      return fail(EHX_ERR_INVALID, "%s: annotation %llu truncated", f.c_str(), (unsigned long long)k);
    }
    memcpy(&width, raw.data() + p, 8);
    p += 8;
    if (width == 0 || width > 8 || (raw.size() - p) / width < rows) {
      delete r;
      return fail(EHX_ERR_INVALID, "%s: annotation %llu of width %llu", f.c_str(), (unsigned long long)k,
                  (unsigned long long)width);
    }
    if (k == 0) {  // the first annotation column is the relation's annotation
      r->annotated = 1;
      r->annos.assign(rows, 0);
      for (uint64_t i = 0; i < rows; i++) memcpy(&r->annos[i], raw.data() + p + i * width, width);
    }
    p += width * rows;
  }
  *out = r;
  return EHX_OK;
}

}  // namespace ehx

int ehx_encoded_open(const char *dir, ehx_refdb **out) { return ehx::encoded_open(dir, out); }

int ehx_trie_build_encoded(ehx_ctx *ctx, const char *dir, int first, int second, int policy, ehx_trie **out) {
  if (!ctx || !out) return ehx::fail(EHX_ERR_INVALID, "null argument");
  ehx_refdb *r = nullptr;
  int rc = ehx::encoded_open(dir, &r);
  if (rc) return rc;
  if (first < 0 || second < 0 || (uint32_t)first >= r->ncols || (uint32_t)second >= r->ncols || first == second) {
    const uint32_t nc = r->ncols;
    delete r;
    return ehx::fail(EHX_ERR_INVALID, "ordering (%d,%d) of a relation with %u attributes", first, second, nc);
  }
  rc = ehx_trie_build(ctx, r->cols[first].data(), r->cols[second].data(), r->rows, policy, 0, out);
  delete r;
  return rc;
}

int ehx_refdb_open(const char *dir, int anno_bytes, ehx_refdb **out) { return ehx::refdb_open(dir, anno_bytes, out); }

int ehx_refdb_info(const ehx_refdb *r, uint32_t *num_columns, uint64_t *rows, int32_t *annotated) {
  if (!r) return ehx::fail(EHX_ERR_INVALID, "null database handle");
  if (num_columns) *num_columns = r->ncols;
  if (rows) *rows = r->rows;
  if (annotated) *annotated = r->annotated;
  return EHX_OK;
}

int ehx_refdb_columns(const ehx_refdb *r, uint32_t *const *cols, uint64_t *annos) {
  if (!r || (r->ncols && !cols)) return ehx::fail(EHX_ERR_INVALID, "null argument");
  for (uint32_t c = 0; c < r->ncols; c++) {
    if (!cols[c]) return ehx::fail(EHX_ERR_INVALID, "null column %u", c);
    if (r->rows) memcpy(cols[c], r->cols[c].data(), 4 * r->rows);
  }
  if (annos && r->annotated && r->rows) memcpy(annos, r->annos.data(), 8 * r->rows);
  return EHX_OK;
}

int ehx_refdb_close(ehx_refdb *r) {
  delete r;
  return EHX_OK;
}

int ehx_trie_load_reference(ehx_ctx *ctx, const char *dir, int anno_bytes, int policy, ehx_trie **out) {
  if (!ctx || !out) return ehx::fail(EHX_ERR_INVALID, "null argument");
  ehx_refdb *r = nullptr;
  int rc = ehx::refdb_open(dir, anno_bytes, &r);
  if (rc) return rc;
  if (r->ncols != 2) {
    const uint32_t nc = r->ncols;
    delete r;
    return ehx::fail(EHX_ERR_UNSUPPORTED, "%s: %u columns; device tries hold binary relations", dir, nc);
  }
  rc = ehx_trie_build(ctx, r->cols[0].data(), r->cols[1].data(), r->rows, policy, 0, out);
  delete r;
  return rc;
}
