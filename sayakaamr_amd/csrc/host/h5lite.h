// h5lite.h -- a small HDF5 file writer with no library dependency.
//
// Writes the subset of the HDF5 file format (format specification 2.0, the layout HDF5 1.8+
// libraries read and write by default) that an AMR output file needs:
//   * superblock version 0 (8-byte offsets and lengths);
//   * groups with a symbol table: version-1 object header, local heap of link names, a version-1
//     B-tree leaf with one symbol-table node holding every link, sorted by name;
//   * scalar attributes (fixed-point, IEEE float, fixed-length string, compound of those);
//   * one-dimensional datasets with contiguous storage, raw data streamed from the caller's
//     buffer after all metadata (no copy).
// Everything little-endian, every structure 8-byte aligned.  The file is written in one pass:
// sizes first, then addresses, then bytes.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace h5lite {

// An encoded datatype message (class, bit fields, size, properties).
struct Type {
    std::vector<uint8_t> enc;
    uint32_t size = 0;   // bytes per element
};

Type int32_type();                 // signed 32-bit little-endian (H5T_NATIVE_INT here)
Type f64_type();                   // IEEE binary64 little-endian (H5T_NATIVE_DOUBLE here)
Type string_type(uint32_t bytes);  // fixed-length, null-terminated, ASCII (H5T_C_S1 sized)

struct Member {
    std::string name;
    uint32_t offset;
    Type type;
};
Type compound_type(uint32_t size, const std::vector<Member>& members);

struct Attribute {  // scalar dataspace
    std::string name;
    Type type;
    std::vector<uint8_t> value;   // type.size bytes
};

struct Dataset {  // rank 1, contiguous
    std::string name;
    Type type;
    uint64_t count = 0;           // elements
    const void* data = nullptr;   // count * type.size bytes, read while the file is written
};

struct Group {
    std::string name;
    std::vector<Attribute> attrs;
    std::vector<Group> groups;
    std::vector<Dataset> datasets;
};

template <class T>
Attribute scalar_attr(const std::string& name, const Type& t, const T& v) {
    Attribute a{name, t, std::vector<uint8_t>(sizeof(T))};
    const uint8_t* p = reinterpret_cast<const uint8_t*>(&v);
    a.value.assign(p, p + sizeof(T));
    return a;
}

// Writes the file rooted at root (root.name is ignored).  Returns 0, or -1 with *err set.
int write(const std::string& path, const Group& root, std::string* err);

}  // namespace h5lite
