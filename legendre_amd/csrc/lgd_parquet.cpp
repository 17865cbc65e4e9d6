// Parquet writer for snapshot columns; see lgd_parquet.h.
#include "lgd_parquet.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <thread>

namespace lgd::pq {
namespace {

using Bytes = std::vector<uint8_t>;

void put_varint(Bytes &o, uint64_t v) {
  while (v >= 0x80) {
    o.push_back((uint8_t)(v | 0x80));
    v >>= 7;
  }
  o.push_back((uint8_t)v);
}

uint64_t zigzag(int64_t v) { return ((uint64_t)v << 1) ^ (uint64_t)(v >> 63); }

// ------------------------------------------------------- Thrift compact protocol
class Thrift {
 public:
  static constexpr uint8_t kTrue = 1, kFalse = 2, kI16 = 4, kI32 = 5, kI64 = 6, kBinary = 8,
                           kList = 9, kStruct = 12;
  explicit Thrift(Bytes &o) : o_(o) {}
  void begin() {
    stack_.push_back(last_);
    last_ = 0;
  }
  void end() {
    o_.push_back(0);  // STOP
    last_ = stack_.back();
    stack_.pop_back();
  }
  void field(int16_t id, uint8_t type) {
    const int d = id - last_;
    if (d > 0 && d <= 15) {
      o_.push_back((uint8_t)((d << 4) | type));
    } else {
      o_.push_back(type);
      put_varint(o_, zigzag(id));
    }
    last_ = id;
  }
  void i16(int16_t id, int16_t v) {
    field(id, kI16);
    put_varint(o_, zigzag(v));
  }
  void i32(int16_t id, int32_t v) {
    field(id, kI32);
    put_varint(o_, zigzag(v));
  }
  void i64(int16_t id, int64_t v) {
    field(id, kI64);
    put_varint(o_, zigzag(v));
  }
  void binary(int16_t id, const void *p, size_t n) {
    field(id, kBinary);
    raw_binary(p, n);
  }
  void string(int16_t id, const std::string &s) { binary(id, s.data(), s.size()); }
  void struct_begin(int16_t id) {
    field(id, kStruct);
    begin();
  }
  void list(int16_t id, uint8_t elem, size_t n) {
    field(id, kList);
    if (n < 15) {
      o_.push_back((uint8_t)((n << 4) | elem));
    } else {
      o_.push_back((uint8_t)(0xF0 | elem));
      put_varint(o_, n);
    }
  }
  void raw_i32(int32_t v) { put_varint(o_, zigzag(v)); }
  void raw_binary(const void *p, size_t n) {
    put_varint(o_, n);
    const uint8_t *b = (const uint8_t *)p;
    o_.insert(o_.end(), b, b + n);
  }

 private:
  Bytes &o_;
  int16_t last_ = 0;
  std::vector<int16_t> stack_;
};

// Parquet enums (parquet.thrift).
constexpr int32_t kPlain = 0, kRle = 3, kRleDictionary = 8;
constexpr int32_t kSnappy = 1;
constexpr int32_t kDataPage = 0, kDictionaryPage = 2;
constexpr int32_t kRequired = 0;

// ---------------------------------------------------------------------- Snappy
inline uint32_t load32(const uint8_t *p) {
  uint32_t v;
  std::memcpy(&v, p, 4);
  return v;
}

void emit_literal(const uint8_t *p, size_t n, Bytes &o) {
  if (n == 0) return;
  const size_t m = n - 1;
  if (m < 60) {
    o.push_back((uint8_t)(m << 2));
  } else {
    int bytes = m < (1u << 8) ? 1 : m < (1u << 16) ? 2 : m < (1u << 24) ? 3 : 4;
    o.push_back((uint8_t)((59 + bytes) << 2));
    for (int i = 0; i < bytes; ++i) o.push_back((uint8_t)(m >> (8 * i)));
  }
  o.insert(o.end(), p, p + n);
}

void emit_copy2(size_t offset, size_t len, Bytes &o) {
  o.push_back((uint8_t)(2 | ((len - 1) << 2)));
  o.push_back((uint8_t)(offset & 0xff));
  o.push_back((uint8_t)(offset >> 8));
}

// A back-reference of `len` >= 4 bytes at distance `offset` (1 <= offset < 65536).
void emit_copy(size_t offset, size_t len, Bytes &o) {
  while (len >= 68) {
    emit_copy2(offset, 64, o);
    len -= 64;
  }
  if (len > 64) {
    emit_copy2(offset, 60, o);
    len -= 60;
  }
  if (len <= 11 && offset < 2048) {
    o.push_back((uint8_t)(1 | ((len - 4) << 2) | ((offset >> 8) << 5)));
    o.push_back((uint8_t)(offset & 0xff));
  } else {
    emit_copy2(offset, len, o);
  }
}

// Greedy LZ77 over one block of at most 64 KiB: a 4-byte hash finds a candidate,
// matches are extended forward, misses step ahead faster through data that does
// not compress.
void compress_block(const uint8_t *src, size_t n, std::vector<uint16_t> &table, Bytes &o) {
  if (n < 8) {
    emit_literal(src, n, o);
    return;
  }
  std::fill(table.begin(), table.end(), (uint16_t)0);
  auto hash = [](uint32_t v) { return (v * 0x1e35a7bdu) >> 18; };
  size_t ip = 1, lit = 0;
  uint32_t skip = 32;
  while (ip + 4 <= n) {
    const uint32_t cur = load32(src + ip);
    const uint32_t h = hash(cur);
    const size_t cand = table[h];
    table[h] = (uint16_t)ip;
    if (cand < ip && load32(src + cand) == cur) {
      size_t len = 4;
      while (ip + len + 8 <= n) {
        uint64_t a, b;
        std::memcpy(&a, src + cand + len, 8);
        std::memcpy(&b, src + ip + len, 8);
        if (a != b) break;
        len += 8;
      }
      while (ip + len < n && src[cand + len] == src[ip + len]) ++len;
      emit_literal(src + lit, ip - lit, o);
      emit_copy(ip - cand, len, o);
      ip += len;
      lit = ip;
      skip = 32;
    } else {
      ip += skip >> 5;
      ++skip;
    }
  }
  emit_literal(src + lit, n - lit, o);
}

// ------------------------------------------------------------ column encoding
// A chunk with more distinct values than this is written PLAIN: Snappy then takes
// the runs, and wide dictionary indices would cost more to pack than they save.
constexpr size_t kMaxDict = 1 << 16;
constexpr size_t kPlainPageValues = 1 << 17;  // 1 MiB PLAIN data pages

struct Encoded {
  Bytes bytes;               // page headers and compressed pages, in file order
  int64_t dict_offset = -1;  // within bytes
  int64_t data_offset = 0;
  int64_t uncompressed = 0;  // headers + uncompressed page bodies
  bool dict = false;
  bool has_stats = false;
  uint64_t min_bits = 0, max_bits = 0;
};

void page(Bytes &out, int32_t type, const uint8_t *body, size_t len, int32_t num_values,
          int32_t encoding, int64_t &uncompressed) {
  thread_local Bytes comp;  // reused across pages
  snappy_compress(body, len, comp);
  Bytes hdr;
  Thrift t(hdr);
  t.begin();
  t.i32(1, type);
  t.i32(2, (int32_t)len);
  t.i32(3, (int32_t)comp.size());
  if (type == kDictionaryPage) {
    t.struct_begin(7);
    t.i32(1, num_values);
    t.i32(2, encoding);
    t.end();
  } else {
    t.struct_begin(5);
    t.i32(1, num_values);
    t.i32(2, encoding);
    t.i32(3, kRle);
    t.i32(4, kRle);
    t.end();
  }
  t.end();
  out.insert(out.end(), hdr.begin(), hdr.end());
  out.insert(out.end(), comp.begin(), comp.end());
  uncompressed += (int64_t)(hdr.size() + len);
}

// The RLE / bit-packed hybrid of `n` indices at `bw` bits: runs of >= 8 equal
// values as RLE runs, everything else bit-packed in groups of 8 (LSB first).
void rle_hybrid(const uint32_t *v, size_t n, int bw, Bytes &o) {
  const size_t vbytes = (size_t)(bw + 7) / 8;
  // Bit-pack v[from, to), zero-padded to whole groups of 8.
  auto pack = [&](size_t from, size_t to) {
    if (to == from) return;
    const size_t groups = (to - from + 7) / 8;
    put_varint(o, (groups << 1) | 1);
    const size_t base = o.size();
    o.resize(base + groups * (size_t)bw, 0);  // 8 values of bw bits = bw bytes
    uint8_t *out = o.data() + base;
    uint64_t acc = 0;
    int nb = 0;
    size_t w = 0;
    for (size_t k = from; k < from + groups * 8; ++k) {
      acc |= (uint64_t)(k < to ? v[k] : 0u) << nb;
      nb += bw;
      while (nb >= 8) {
        out[w++] = (uint8_t)acc;
        acc >>= 8;
        nb -= 8;
      }
    }
  };
  size_t i = 0, lit = 0;  // v[lit, i) waits to be bit-packed
  while (i < n) {
    size_t j = i + 1;
    while (j < n && v[j] == v[i]) ++j;
    size_t run = j - i;
    if (run >= 8) {
      const size_t top = (8 - (i - lit) % 8) % 8;  // complete the pending group
      i += top;
      run -= top;
      if (run >= 8) {
        pack(lit, i);
        put_varint(o, (uint64_t)run << 1);
        for (size_t k = 0; k < vbytes; ++k) o.push_back((uint8_t)(v[i] >> (8 * k)));
        i += run;
        lit = i;
      }
      continue;
    }
    i = j;
    if (i - lit >= 4096 && (i - lit) % 8 == 0) {
      pack(lit, i);
      lit = i;
    }
  }
  pack(lit, n);
}

int bit_width(size_t max_index) {
  int b = 0;
  while (max_index >> b) ++b;
  return b;
}

// Open-addressing dictionary of 64-bit patterns (values compare by bits, so -0.0,
// +0.0 and distinct NaNs stay distinct and round-trip exactly).
struct Dict {
  std::vector<uint64_t> keys;
  std::vector<int32_t> slot;  // -1 = empty
  std::vector<uint64_t> values;
  size_t mask;
  Dict() : keys(kMaxDict * 2), slot(kMaxDict * 2, -1), mask(kMaxDict * 2 - 1) {}
  // Index of `k`, inserting it; -1 when the dictionary is full.
  int32_t index(uint64_t k) {
    size_t h = (size_t)((k * 0x9E3779B97F4A7C15ull) >> 40) & mask;
    while (slot[h] >= 0) {
      if (keys[h] == k) return slot[h];
      h = (h + 1) & mask;
    }
    if (values.size() >= kMaxDict) return -1;
    keys[h] = k;
    slot[h] = (int32_t)values.size();
    values.push_back(k);
    return slot[h];
  }
};

void update_stats(Type type, uint64_t bits, Encoded &e) {
  if (type == Type::kDouble) {
    double v;
    std::memcpy(&v, &bits, 8);
    if (std::isnan(v)) return;
    double lo, hi;
    std::memcpy(&lo, &e.min_bits, 8);
    std::memcpy(&hi, &e.max_bits, 8);
    if (!e.has_stats || v < lo || (v == lo && std::signbit(v))) e.min_bits = bits;
    if (!e.has_stats || v > hi || (v == hi && !std::signbit(v))) e.max_bits = bits;
  } else {
    const int64_t v = (int64_t)bits;
    if (!e.has_stats || v < (int64_t)e.min_bits) e.min_bits = bits;
    if (!e.has_stats || v > (int64_t)e.max_bits) e.max_bits = bits;
  }
  e.has_stats = true;
}

// Min / max of n values (NaN skipped), merged into e. Zero signs are settled when
// the statistics are written.
void range_stats(Type type, const uint8_t *p, size_t n, Encoded &e) {
  if (type == Type::kDouble) {
    double lo = HUGE_VAL, hi = -HUGE_VAL;
    bool any = false;
    for (size_t i = 0; i < n; ++i) {
      double x;
      std::memcpy(&x, p + 8 * i, 8);
      if (x != x) continue;
      lo = x < lo ? x : lo;
      hi = x > hi ? x : hi;
      any = true;
    }
    if (!any) return;
    uint64_t lb, hb;
    std::memcpy(&lb, &lo, 8);
    std::memcpy(&hb, &hi, 8);
    update_stats(type, lb, e);
    update_stats(type, hb, e);
  } else {
    int64_t lo = INT64_MAX, hi = INT64_MIN;
    for (size_t i = 0; i < n; ++i) {
      int64_t x;
      std::memcpy(&x, p + 8 * i, 8);
      lo = x < lo ? x : lo;
      hi = x > hi ? x : hi;
    }
    if (n == 0) return;
    update_stats(type, (uint64_t)lo, e);
    update_stats(type, (uint64_t)hi, e);
  }
}

// One column chunk: rows [r0, r0 + n) of `c`.
void encode_chunk(const Column &c, int64_t r0, int64_t n, Encoded &e) {
  auto value = [&](int64_t r) -> uint64_t {
    if (!c.data) return c.constant;
    uint64_t b;
    std::memcpy(&b, (const uint8_t *)c.data + 8 * (r0 + r), 8);
    return b;
  };
  e.bytes.reserve((size_t)n * 8 / 2 + 4096);  // address space only until written
  std::vector<uint32_t> idx;
  Dict d;
  bool dict = true;
  if (!c.data) {
    d.index(c.constant);  // one dictionary entry, every index 0
    idx.assign((size_t)n, 0u);
  } else {
    uint64_t prev = 0;
    int32_t prev_k = -1;
    for (int64_t r = 0; r < n; ++r) {
      const uint64_t v = value(r);
      if (prev_k >= 0 && v == prev) {  // runs of one value skip the hash
        idx.push_back((uint32_t)prev_k);
        continue;
      }
      const int32_t k = d.index(v);
      if (k < 0) {
        dict = false;
        break;
      }
      idx.push_back((uint32_t)k);
      prev = v;
      prev_k = k;
    }
  }
  e.dict = dict;
  if (dict) {
    range_stats(c.type, (const uint8_t *)d.values.data(), d.values.size(), e);
    e.dict_offset = 0;
    page(e.bytes, kDictionaryPage, (const uint8_t *)d.values.data(), d.values.size() * 8,
         (int32_t)d.values.size(), kPlain, e.uncompressed);
    e.data_offset = (int64_t)e.bytes.size();
    const int bw = bit_width(d.values.empty() ? 0 : d.values.size() - 1);
    size_t per_page = std::max<size_t>(8192, ((size_t)1 << 23) / (size_t)std::max(bw, 1));
    per_page &= ~(size_t)7;
    for (int64_t p = 0; p < n; p += (int64_t)per_page) {
      const size_t m = (size_t)std::min<int64_t>((int64_t)per_page, n - p);
      Bytes b;
      b.push_back((uint8_t)bw);
      rle_hybrid(idx.data() + p, m, bw, b);
      page(e.bytes, kDataPage, b.data(), b.size(), (int32_t)m, kRleDictionary, e.uncompressed);
    }
  } else {
    // PLAIN pages straight from the column (a constant column always takes the
    // dictionary path, so data is set here).
    e.data_offset = 0;
    const uint8_t *src = (const uint8_t *)c.data + 8 * r0;
    range_stats(c.type, src, (size_t)n, e);
    for (int64_t p = 0; p < n; p += (int64_t)kPlainPageValues) {
      const size_t m = (size_t)std::min<int64_t>((int64_t)kPlainPageValues, n - p);
      page(e.bytes, kDataPage, src + 8 * p, m * 8, (int32_t)m, kPlain, e.uncompressed);
    }
  }
}

// Parquet's statistics convention for a zero bound of a float column: a minimum of
// +0 is written as -0 and a maximum of -0 as +0.
void write_stats(Thrift &t, const Column &c, const Encoded &e) {
  uint64_t lo = e.min_bits, hi = e.max_bits;
  if (c.type == Type::kDouble) {
    if (lo == 0) lo = 0x8000000000000000ull;
    if (hi == 0x8000000000000000ull) hi = 0;
  }
  t.struct_begin(12);
  t.i64(3, 0);  // null_count
  t.binary(5, &hi, 8);
  t.binary(6, &lo, 8);
  t.end();
}

}  // namespace

void snappy_compress(const uint8_t *in, size_t n, std::vector<uint8_t> &out) {
  out.clear();
  out.reserve(n + n / 6 + 32);
  put_varint(out, n);
  std::vector<uint16_t> table(1 << 14);
  for (size_t base = 0; base < n; base += (size_t)1 << 16)
    compress_block(in + base, std::min((size_t)1 << 16, n - base), table, out);
}

Status write_file(const std::string &path, const std::vector<Column> &cols, int64_t nrows,
                  const std::vector<int64_t> &group_rows, std::string *err) {
  auto bad = [&](Status s, const std::string &m) {
    if (err) *err = m;
    return s;
  };
  if (cols.empty()) return bad(Status::kInvalid, "no columns");
  int64_t total = 0;
  for (int64_t g : group_rows) {
    if (g < 0) return bad(Status::kInvalid, "negative row group");
    total += g;
  }
  if (total != nrows) return bad(Status::kInvalid, "row groups do not sum to the row count");
  const std::string tmp = path + ".tmp";
  FILE *f = std::fopen(tmp.c_str(), "wb");
  if (!f) return bad(Status::kIo, "cannot create " + tmp);
  int64_t pos = 0;
  bool ok = true;
  auto put = [&](const void *p, size_t n) {
    if (ok && n && std::fwrite(p, 1, n, f) != n) ok = false;
    pos += (int64_t)n;
  };
  put("PAR1", 4);

  struct ChunkMeta {
    int64_t offset, data_offset, dict_offset, uncompressed, compressed;
    bool dict, has_stats;
    uint64_t lo, hi;
  };
  struct GroupMeta {
    int64_t rows, offset;
    std::vector<ChunkMeta> chunks;
  };
  std::vector<GroupMeta> groups;
  const size_t ncol = cols.size();
  const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<int64_t> gr, g0;  // non-empty row groups and their first rows
  for (int64_t rows : group_rows) {
    if (rows > 0) {
      g0.push_back(gr.empty() ? 0 : g0.back() + gr.back());
      gr.push_back(rows);
    }
  }
  // Column chunks are encoded in parallel over a window of row groups (bounding the
  // encoded bytes held at once), then written in file order.
  const size_t window = std::max<size_t>(1, hw / ncol);
  for (size_t w0 = 0; w0 < gr.size(); w0 += window) {
    const size_t w1 = std::min(gr.size(), w0 + window);
    const size_t jobs = (w1 - w0) * ncol;
    std::vector<Encoded> enc(jobs);
    std::vector<std::thread> pool;
    const unsigned nt = (unsigned)std::min<size_t>(hw, jobs);
    for (unsigned t0 = 0; t0 < nt; ++t0)
      pool.emplace_back([&, t0]() {
        for (size_t j = t0; j < jobs; j += nt) {
          const size_t g = w0 + j / ncol, c = j % ncol;
          encode_chunk(cols[c], g0[g], gr[g], enc[j]);
        }
      });
    for (auto &th : pool) th.join();
    for (size_t g = w0; g < w1; ++g) {
      GroupMeta gm{gr[g], pos, {}};
      for (size_t c = 0; c < ncol; ++c) {
        const Encoded &e = enc[(g - w0) * ncol + c];
        ChunkMeta m;
        m.offset = pos;
        m.dict = e.dict;
        m.dict_offset = e.dict ? pos + e.dict_offset : -1;
        m.data_offset = pos + e.data_offset;
        m.uncompressed = e.uncompressed;
        m.compressed = (int64_t)e.bytes.size();
        m.has_stats = e.has_stats;
        m.lo = e.min_bits;
        m.hi = e.max_bits;
        put(e.bytes.data(), e.bytes.size());
        gm.chunks.push_back(m);
      }
      groups.push_back(gm);
    }
  }

  // FileMetaData.
  Bytes meta;
  Thrift t(meta);
  t.begin();
  t.i32(1, 1);  // version
  t.list(2, Thrift::kStruct, ncol + 1);
  t.begin();
  t.string(4, "snapshot");
  t.i32(5, (int32_t)ncol);
  t.end();
  for (const Column &c : cols) {
    t.begin();
    t.i32(1, (int32_t)c.type);
    t.i32(3, kRequired);
    t.string(4, c.name);
    t.end();
  }
  t.i64(3, nrows);
  t.list(4, Thrift::kStruct, groups.size());
  for (size_t g = 0; g < groups.size(); ++g) {
    const GroupMeta &gm = groups[g];
    t.begin();
    t.list(1, Thrift::kStruct, ncol);
    int64_t bytes = 0, comp = 0;
    for (size_t c = 0; c < ncol; ++c) {
      const ChunkMeta &m = gm.chunks[c];
      bytes += m.uncompressed;
      comp += m.compressed;
      t.begin();
      t.i64(2, m.offset);
      t.struct_begin(3);
      t.i32(1, (int32_t)cols[c].type);
      if (m.dict) {
        t.list(2, Thrift::kI32, 2);
        t.raw_i32(kPlain);
        t.raw_i32(kRleDictionary);
      } else {
        t.list(2, Thrift::kI32, 1);
        t.raw_i32(kPlain);
      }
      t.list(3, Thrift::kBinary, 1);
      t.raw_binary(cols[c].name.data(), cols[c].name.size());
      t.i32(4, kSnappy);
      t.i64(5, gm.rows);
      t.i64(6, m.uncompressed);
      t.i64(7, m.compressed);
      t.i64(9, m.data_offset);
      if (m.dict) t.i64(11, m.dict_offset);
      if (m.has_stats) {
        Encoded e;
        e.min_bits = m.lo;
        e.max_bits = m.hi;
        write_stats(t, cols[c], e);
      }
      t.end();  // ColumnMetaData
      t.end();  // ColumnChunk
    }
    t.i64(2, bytes);
    t.i64(3, gm.rows);
    t.i64(5, gm.offset);
    t.i64(6, comp);
    t.i16(7, (int16_t)g);
    t.end();
  }
  t.string(6, "legendre_amd parquet writer");
  // column_orders: TYPE_ORDER for every column, so readers use min/max_value.
  t.list(7, Thrift::kStruct, ncol);
  for (size_t c = 0; c < ncol; ++c) {
    t.begin();
    t.struct_begin(1);
    t.end();
    t.end();
  }
  t.end();
  put(meta.data(), meta.size());
  const uint32_t len = (uint32_t)meta.size();
  uint8_t tail[8];
  std::memcpy(tail, &len, 4);
  std::memcpy(tail + 4, "PAR1", 4);
  put(tail, 8);
  if (std::fclose(f) != 0) ok = false;
  if (!ok || std::rename(tmp.c_str(), path.c_str()) != 0) {
    std::remove(tmp.c_str());
    return bad(Status::kIo, "write failed: " + path);
  }
  return Status::kOk;
}

}  // namespace lgd::pq

extern "C" int lgd_parquet_write_file(const char *path, size_t ncols, const char *const *names,
                                      const int32_t *types, const void *const *columns,
                                      const uint64_t *constants, int64_t nrows,
                                      const int64_t *group_rows, size_t ngroups, char *err,
                                      size_t errlen) {
  constexpr int kOk = 0, kInvalid = 1, kIo = 8;  // lgd_status
  auto msg = [&](const std::string &m) {
    if (err && errlen) std::snprintf(err, errlen, "%s", m.c_str());
  };
  if (!path || !names || !types || !columns || (ngroups && !group_rows) || nrows < 0) {
    msg("null argument");
    return kInvalid;
  }
  std::vector<lgd::pq::Column> cols(ncols);
  for (size_t c = 0; c < ncols; ++c) {
    if (!names[c] || (types[c] != 0 && types[c] != 1) || (!columns[c] && !constants)) {
      msg("invalid column");
      return kInvalid;
    }
    cols[c].name = names[c];
    cols[c].type = types[c] == 0 ? lgd::pq::Type::kDouble : lgd::pq::Type::kInt64;
    cols[c].data = columns[c];
    cols[c].constant = columns[c] ? 0 : constants[c];
  }
  std::string e;
  const lgd::pq::Status st = lgd::pq::write_file(
      path, cols, nrows, std::vector<int64_t>(group_rows, group_rows + ngroups), &e);
  msg(e);
  return st == lgd::pq::Status::kOk ? kOk : st == lgd::pq::Status::kIo ? kIo : kInvalid;
}
