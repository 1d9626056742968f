// zj_h5.hpp -- a small, dependency-free HDF5 codec for flat files: scalar
// attributes and one-dimensional datasets in the root group.
//
// The restart files of the reference (src/io/restart.cpp:8-158) are written
// by libhdf5 with default property lists: file format 0 (superblock v0,
// version-1 object headers, root group as a symbol table: v1 B-tree + local
// heap + symbol-table nodes), scalar attributes on the root group, contiguous
// datasets.  This codec writes exactly that structure from the published
// HDF5 File Format Specification, and reads it back whether libhdf5 or this
// writer produced it (continuation chunks, NIL messages, modification-time and
// fill-value messages, attribute messages v1-v3, dataspace v1/v2, layout
// v3/v4 contiguous or compact).  Anything else (chunked or filtered storage,
// shared datatypes, superblock v2/v3) is refused with a message.
#pragma once

#include <cstdint>
#include <map>
#include <string>
#include <vector>

namespace zj {
namespace h5 {

enum class Type : int { F64 = 0, I32 = 1, I64 = 2, F32 = 3, STR = 4 };

struct Attribute {
    std::string name;
    Type type = Type::F64;
    double f = 0.0;       // F64 / F32
    int64_t i = 0;        // I32 / I64
    std::string s;        // STR (fixed-length, NUL-terminated)
};

struct DatasetOut {
    std::string name;
    Type type = Type::F64;   // F64, F32, I32 or I64
    const void* data = nullptr;
    uint64_t count = 0;      // elements (1-D)
};

// Creates (truncates) `path`.  Attributes are stored on the root group in the
// order given; datasets are contiguous, little-endian.  Returns false and sets
// err on failure.
bool write_file(const std::string& path, const std::vector<Attribute>& attrs,
                const std::vector<DatasetOut>& dsets, std::string& err);

struct DatasetInfo {
    Type type = Type::F64;
    uint64_t count = 0;       // product of the dimensions
    int rank = 0;
    bool compact = false;
    uint64_t addr = 0;        // contiguous: absolute file offset of the data
    uint64_t size = 0;        // bytes of stored data
    std::vector<uint8_t> compact_data;
};

class Reader {
public:
    Reader() = default;
    ~Reader();
    Reader(const Reader&) = delete;
    Reader& operator=(const Reader&) = delete;

    bool open(const std::string& path, std::string& err);
    bool hasAttribute(const std::string& name) const { return attrs_.count(name) != 0; }
    bool hasDataset(const std::string& name) const { return links_.count(name) != 0; }
    bool attribute(const std::string& name, double& v, std::string& err) const;
    bool attribute(const std::string& name, int64_t& v, std::string& err) const;
    bool attribute(const std::string& name, std::string& v, std::string& err) const;
    bool dataset(const std::string& name, DatasetInfo& info, std::string& err) const;
    // Reads and converts all `count` elements (count must equal the stored
    // element count).  Integers convert to double exactly below 2^53; reading
    // an integer dataset as int32 checks the range.
    bool read(const std::string& name, double* out, uint64_t count, std::string& err) const;
    bool read(const std::string& name, int32_t* out, uint64_t count, std::string& err) const;
    std::vector<std::string> datasetNames() const;
    std::vector<std::string> attributeNames() const;

private:
    struct RawAttr {
        Type type;
        bool is_signed;
        int size;
        std::vector<uint8_t> data;
    };
    bool pread_at(uint64_t addr, void* dst, uint64_t n, std::string& err) const;
    int fd_ = -1;
    uint64_t file_size_ = 0;
    uint64_t base_ = 0;
    std::map<std::string, RawAttr> attrs_;
    std::map<std::string, uint64_t> links_;   // name -> object header address

    friend struct ReaderImpl;
};

}  // namespace h5
}  // namespace zj
