// h5lite.cpp -- see h5lite.h.  Structure layouts follow the HDF5 file format specification
// (superblock v0, object header v1, local heap v0, v1 B-tree node type 0, symbol table node v1,
// dataspace v1, datatype v1, data layout v3, attribute v1).
#include "h5lite.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>

namespace h5lite {

namespace {

constexpr uint64_t UNDEF = ~0ull;
constexpr uint32_t BTREE_K = 16;       // group B-tree: up to 2K children per node
constexpr uint64_t HEAP_FREE_NONE = 1;  // free-list head of a local heap without free space
constexpr size_t SUPERBLOCK = 96;
constexpr size_t ENTRY = 40;           // symbol table entry
constexpr size_t OH_PREFIX = 16;       // object header v1 prefix (12 bytes, aligned to 8)

enum MsgType : uint16_t { M_DATASPACE = 1, M_DATATYPE = 3, M_LAYOUT = 8, M_ATTRIBUTE = 12, M_STAB = 17 };

size_t pad8(size_t n) { return (n + 7) & ~size_t(7); }

struct Out {
    std::vector<uint8_t> b;
    void u8(uint8_t v) { b.push_back(v); }
    void u16(uint16_t v) { for (int i = 0; i < 2; i++) b.push_back(uint8_t(v >> (8 * i))); }
    void u32(uint32_t v) { for (int i = 0; i < 4; i++) b.push_back(uint8_t(v >> (8 * i))); }
    void u64(uint64_t v) { for (int i = 0; i < 8; i++) b.push_back(uint8_t(v >> (8 * i))); }
    void raw(const void* p, size_t n) {
        const uint8_t* q = static_cast<const uint8_t*>(p);
        b.insert(b.end(), q, q + n);
    }
    void zeros(size_t n) { b.insert(b.end(), n, 0); }
    void align8() { b.resize(pad8(b.size()), 0); }
    void name(const std::string& s) {   // null-terminated, length padded to a multiple of 8
        raw(s.data(), s.size());
        zeros(pad8(s.size() + 1) - s.size());
    }
};

Type make_type(uint8_t cls, uint32_t bits, uint32_t size) {
    Type t;
    t.size = size;
    Out o;
    o.u8(uint8_t(cls | (1 << 4)));   // version 1
    o.u8(uint8_t(bits));
    o.u8(uint8_t(bits >> 8));
    o.u8(uint8_t(bits >> 16));
    o.u32(size);
    t.enc = std::move(o.b);
    return t;
}

// one header message: type, data (padded), flags
struct Msg {
    uint16_t type;
    uint8_t flags;
    std::vector<uint8_t> data;
};

std::vector<uint8_t> dataspace_scalar() {
    Out o;
    o.u8(1);   // version
    o.u8(0);   // rank 0: scalar
    o.u8(0);   // flags: no maximum dimensions
    o.zeros(5);
    return o.b;
}

std::vector<uint8_t> dataspace_1d(uint64_t n) {
    Out o;
    o.u8(1);
    o.u8(1);
    o.u8(0);
    o.zeros(5);
    o.u64(n);
    return o.b;
}

Msg attr_msg(const Attribute& a) {
    const std::vector<uint8_t> sp = dataspace_scalar();
    Out o;
    o.u8(1);   // version
    o.u8(0);
    o.u16(uint16_t(a.name.size() + 1));
    o.u16(uint16_t(a.type.enc.size()));
    o.u16(uint16_t(sp.size()));
    o.name(a.name);
    o.raw(a.type.enc.data(), a.type.enc.size());
    o.align8();
    o.raw(sp.data(), sp.size());
    o.align8();
    o.raw(a.value.data(), a.value.size());
    return Msg{M_ATTRIBUTE, 0, std::move(o.b)};
}

size_t header_size(const std::vector<Msg>& m) {
    size_t s = OH_PREFIX;
    for (const Msg& x : m) s += 8 + pad8(x.data.size());
    return s;
}

void emit_header(Out& o, const std::vector<Msg>& m) {
    o.u8(1);   // version
    o.u8(0);
    o.u16(uint16_t(m.size()));
    o.u32(1);   // reference count
    o.u32(uint32_t(header_size(m) - OH_PREFIX));
    o.u32(0);
    for (const Msg& x : m) {
        o.u16(x.type);
        o.u16(uint16_t(pad8(x.data.size())));
        o.u8(x.flags);
        o.zeros(3);
        o.raw(x.data.data(), x.data.size());
        o.align8();
    }
}

struct Link {
    std::string name;
    bool is_group;
    size_t index;      // into the plan's groups or datasets
    uint64_t heap_off;
};

struct DsetPlan {
    const Dataset* d;
    std::vector<Msg> msgs;   // layout message patched once the raw address is known
    uint64_t oh = 0, raw = UNDEF, bytes = 0;
};

struct GroupPlan {
    const Group* g;
    std::vector<Link> links;        // sorted by name (byte order)
    std::vector<size_t> sub;        // GroupPlan indices of the child groups, in g->groups order
    std::vector<size_t> dsets;      // DsetPlan indices, in g->datasets order
    std::vector<Msg> msgs;          // symbol table message first (patched), then attributes
    uint64_t oh = 0, heap = 0, heap_data = 0, btree = 0, snod = UNDEF;
    size_t heap_size = 0;
};

struct Plan {
    std::vector<GroupPlan> groups;
    std::vector<DsetPlan> dsets;
    uint32_t leaf_k = 4;
    size_t btree_size() const { return 24 + 2 * BTREE_K * 8 + (2 * BTREE_K + 1) * 8; }
    size_t snod_size() const { return 8 + 2 * size_t(leaf_k) * ENTRY; }
};

size_t plan_group(Plan& P, const Group& g, std::string* err) {
    const size_t id = P.groups.size();
    P.groups.push_back(GroupPlan{&g, {}, {}, {}, {}});
    std::vector<Link> links;
    for (size_t i = 0; i < g.groups.size(); i++) links.push_back(Link{g.groups[i].name, true, i, 0});
    for (size_t i = 0; i < g.datasets.size(); i++) links.push_back(Link{g.datasets[i].name, false, i, 0});
    std::sort(links.begin(), links.end(), [](const Link& a, const Link& b) { return a.name < b.name; });
    for (size_t i = 0; i < links.size(); i++) {
        if (links[i].name.empty() || links[i].name.find('/') != std::string::npos) {
            *err = "bad link name '" + links[i].name + "'";
            return SIZE_MAX;
        }
        if (i > 0 && links[i].name == links[i - 1].name) {
            *err = "duplicate link name '" + links[i].name + "'";
            return SIZE_MAX;
        }
    }
    // local heap: "" at offset 0, then every link name
    size_t off = 8;
    for (Link& l : links) {
        l.heap_off = off;
        off += pad8(l.name.size() + 1);
    }
    P.leaf_k = std::max<uint32_t>(P.leaf_k, uint32_t((links.size() + 1) / 2));
    {
        GroupPlan& gp = P.groups[id];
        gp.links = links;
        gp.heap_size = off;
        gp.msgs.push_back(Msg{M_STAB, 0, std::vector<uint8_t>(16, 0)});
        for (const Attribute& a : g.attrs) {
            if (a.value.size() != a.type.size) {
                *err = "attribute '" + a.name + "': value size differs from its type";
                return SIZE_MAX;
            }
            gp.msgs.push_back(attr_msg(a));
        }
    }
    for (const Group& c : g.groups) {
        const size_t s = plan_group(P, c, err);
        if (s == SIZE_MAX) return s;
        P.groups[id].sub.push_back(s);
    }
    for (const Dataset& d : g.datasets) {
        DsetPlan dp{&d, {}};
        dp.bytes = d.count * d.type.size;
        if (dp.bytes > 0 && !d.data) {
            *err = "dataset '" + d.name + "' without data";
            return SIZE_MAX;
        }
        dp.msgs.push_back(Msg{M_DATASPACE, 0, dataspace_1d(d.count)});
        dp.msgs.push_back(Msg{M_DATATYPE, 1, d.type.enc});   // constant
        dp.msgs.push_back(Msg{M_LAYOUT, 0, std::vector<uint8_t>(18, 0)});
        P.groups[id].dsets.push_back(P.dsets.size());
        P.dsets.push_back(std::move(dp));
    }
    return id;
}

void assign(Plan& P, uint64_t& at) {
    for (GroupPlan& g : P.groups) {
        g.oh = at;
        at += header_size(g.msgs);
        g.heap = at;
        at += 32;
        g.heap_data = at;
        at += g.heap_size;
        g.btree = at;
        at += P.btree_size();
        if (!g.links.empty()) {
            g.snod = at;
            at += P.snod_size();
        }
    }
    for (DsetPlan& d : P.dsets) {
        d.oh = at;
        at += header_size(d.msgs);
    }
    for (DsetPlan& d : P.dsets) {
        if (d.bytes == 0) continue;
        d.raw = at;
        at += pad8(d.bytes);
    }
}

void entry(Out& o, uint64_t name_off, uint64_t oh, const GroupPlan* g) {
    o.u64(name_off);
    o.u64(oh);
    o.u32(g ? 1 : 0);   // cache type 1: B-tree and heap addresses in the scratch pad
    o.u32(0);
    o.u64(g ? g->btree : 0);
    o.u64(g ? g->heap : 0);
}

}  // namespace

Type int32_type() {
    Type t = make_type(0, 0x08, 4);   // fixed-point, little-endian, signed
    Out o;
    o.u16(0);    // bit offset
    o.u16(32);   // precision
    t.enc.insert(t.enc.end(), o.b.begin(), o.b.end());
    return t;
}

Type f64_type() {
    // floating-point, little-endian, implied leading mantissa bit, sign at bit 63
    Type t = make_type(1, 0x20 | (63u << 8), 8);
    Out o;
    o.u16(0);    // bit offset
    o.u16(64);   // precision
    o.u8(52);    // exponent location
    o.u8(11);    // exponent size
    o.u8(0);     // mantissa location
    o.u8(52);    // mantissa size
    o.u32(1023);
    t.enc.insert(t.enc.end(), o.b.begin(), o.b.end());
    return t;
}

Type string_type(uint32_t bytes) { return make_type(3, 0, bytes); }   // null-terminated, ASCII

Type compound_type(uint32_t size, const std::vector<Member>& members) {
    Type t = make_type(6, uint32_t(members.size()), size);
    Out o;
    for (const Member& m : members) {
        o.name(m.name);
        o.u32(m.offset);
        o.u8(0);       // dimensionality (scalar member)
        o.zeros(3);
        o.u32(0);      // dimension permutation
        o.u32(0);
        o.zeros(16);   // dimension sizes
        o.raw(m.type.enc.data(), m.type.enc.size());
    }
    t.enc.insert(t.enc.end(), o.b.begin(), o.b.end());
    return t;
}

int write(const std::string& path, const Group& root, std::string* err) {
    std::string e;
    Plan P;
    if (plan_group(P, root, &e) == SIZE_MAX) {
        if (err) *err = e;
        return -1;
    }
    uint64_t eof = SUPERBLOCK;
    assign(P, eof);
    // patch the addresses into the symbol table and layout messages
    for (GroupPlan& g : P.groups) {
        Out o;
        o.u64(g.btree);
        o.u64(g.heap);
        g.msgs[0].data = o.b;
    }
    for (DsetPlan& d : P.dsets) {
        Out o;
        o.u8(3);   // layout version 3
        o.u8(1);   // contiguous
        o.u64(d.raw);
        o.u64(d.bytes);
        d.msgs[2].data = o.b;
    }
    Out o;
    o.raw("\x89HDF\r\n\x1a\n", 8);
    o.u8(0);   // superblock version
    o.u8(0);   // free-space version
    o.u8(0);   // root group symbol table entry version
    o.u8(0);
    o.u8(0);   // shared header message format version
    o.u8(8);   // size of offsets
    o.u8(8);   // size of lengths
    o.u8(0);
    o.u16(uint16_t(P.leaf_k));
    o.u16(uint16_t(BTREE_K));
    o.u32(0);       // file consistency flags
    o.u64(0);       // base address
    o.u64(UNDEF);   // free-space info
    o.u64(eof);
    o.u64(UNDEF);   // driver information block
    entry(o, 0, P.groups[0].oh, &P.groups[0]);
    for (const GroupPlan& g : P.groups) {
        emit_header(o, g.msgs);
        o.raw("HEAP", 4);
        o.u8(0);
        o.zeros(3);
        o.u64(g.heap_size);
        o.u64(HEAP_FREE_NONE);
        o.u64(g.heap_data);
        o.zeros(8);
        for (const Link& l : g.links) o.name(l.name);
        const size_t bt = o.b.size();
        o.raw("TREE", 4);
        o.u8(0);   // group node
        o.u8(0);   // leaf level
        o.u16(g.links.empty() ? 0 : 1);
        o.u64(UNDEF);
        o.u64(UNDEF);
        o.u64(0);   // key 0: ""
        if (!g.links.empty()) {
            o.u64(g.snod);
            o.u64(g.links.back().heap_off);   // key 1: the largest name in the child
        }
        o.zeros(P.btree_size() - (o.b.size() - bt));
        if (!g.links.empty()) {
            const size_t sn = o.b.size();
            o.raw("SNOD", 4);
            o.u8(1);
            o.u8(0);
            o.u16(uint16_t(g.links.size()));
            for (const Link& l : g.links) {
                if (l.is_group) {
                    const GroupPlan& c = P.groups[g.sub[l.index]];
                    entry(o, l.heap_off, c.oh, &c);
                } else {
                    entry(o, l.heap_off, P.dsets[g.dsets[l.index]].oh, nullptr);
                }
            }
            o.zeros(P.snod_size() - (o.b.size() - sn));
        }
    }
    for (const DsetPlan& d : P.dsets) emit_header(o, d.msgs);
    std::unique_ptr<FILE, int (*)(FILE*)> f(std::fopen(path.c_str(), "wb"), &std::fclose);
    if (!f) {
        if (err) *err = "cannot open '" + path + "' for writing";
        return -1;
    }
    bool ok = std::fwrite(o.b.data(), 1, o.b.size(), f.get()) == o.b.size();
    uint64_t at = o.b.size();
    static const uint8_t zero[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (const DsetPlan& d : P.dsets) {
        if (!ok || d.bytes == 0) continue;
        if (at != d.raw) ok = false;
        ok = ok && std::fwrite(d.d->data, 1, d.bytes, f.get()) == d.bytes;
        const size_t padn = pad8(d.bytes) - d.bytes;
        ok = ok && std::fwrite(zero, 1, padn, f.get()) == padn;
        at += pad8(d.bytes);
    }
    ok = ok && at == eof;
    if (std::fclose(f.release()) != 0) ok = false;
    if (!ok) {
        if (err) *err = "write to '" + path + "' failed";
        return -1;
    }
    return 0;
}

}  // namespace h5lite
