// Host-side Parquet file writer for snapshot columns (the layout of the reference's
// ParquetObserver, src/io/parquet.rs): one flat message of required DOUBLE / INT64
// columns, row groups of given sizes, snappy-compressed v1 data pages, a dictionary
// page wherever a column chunk has few distinct values, min/max statistics.
//
// Format sources: the Apache Parquet format specification (file layout, Thrift
// metadata, PLAIN / RLE_DICTIONARY encodings, the RLE/bit-packed hybrid) and the
// Thrift compact protocol specification; the Snappy block format description.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace lgd::pq {

enum class Type : int32_t { kInt64 = 2, kDouble = 5 };

struct Column {
  std::string name;
  Type type = Type::kDouble;
  // nrows values (f64 or i64 by type), or nullptr with `constant` holding the bits
  // of a value every row shares.
  const void *data = nullptr;
  uint64_t constant = 0;
};

enum class Status { kOk = 0, kInvalid = 1, kIo = 2 };

// Writes `cols` (each nrows long) to `path` as row groups of group_rows[i] rows
// (summing to nrows), through a temporary file renamed into place, so a reader never
// sees a partial file. On failure sets *err and removes the temporary file.
Status write_file(const std::string &path, const std::vector<Column> &cols, int64_t nrows,
                  const std::vector<int64_t> &group_rows, std::string *err);

// Snappy block compression of n bytes (raw format: length preamble + elements).
void snappy_compress(const uint8_t *in, size_t n, std::vector<uint8_t> &out);

}  // namespace lgd::pq
