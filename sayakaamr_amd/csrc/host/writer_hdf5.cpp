// writer_hdf5.cpp -- Chombo-format HDF5 output (WriteDataHdf5, SayakaWriterHdf5.cpp:17-530)
// through the build's own HDF5 writer (h5lite.h).
//
// File contents, as the reference writes them (SayakaWriterHdf5.cpp:80-349):
//   /  attributes num_components, num_levels, iteration (= meta.step), time, aspect_ratio
//      (compound x, y[, z] doubles: 1, dy/dx[, dz/dx] of block 0), data_centering (0: cell),
//      prob_lo (compound), component_<i> (64-byte string, meta.name[i])
//   /Chombo_global  attributes SpaceDim, testReal (0.0)
//   /level_<l>, l = 0 .. finest-1 (tree level l+1)  attributes ref_ratio (2), dt, dx (x cell
//      size: block 0's halved l times), prob_domain (compound lo_i, lo_j[, lo_k], hi_i, hi_j
//      [, hi_k] ints); datasets boxes (one such compound per block of the level, ascending id)
//      and data:datatype=0 (doubles: per block, per component, its valid cells, i fastest);
//      group data_attributes with attribute comps (= num_components).
// The reference creates both datasets and then writes one hyperslab per block and component;
// here each dataset is written once from a buffer already in that order.
#include <algorithm>
#include <cstring>

#include "h5lite.h"
#include "sayaka_host.h"
#include "writer.h"

namespace sayaka {
inline namespace amd {

namespace {

const char* AXES[3] = {"x", "y", "z"};
const char* LO[3] = {"lo_i", "lo_j", "lo_k"};
const char* HI[3] = {"hi_i", "hi_j", "hi_k"};

// realvect_t (SayakaWriterHdf5.cpp:20-26): nd doubles
h5lite::Type realvect_type(int nd) {
    std::vector<h5lite::Member> m;
    for (int d = 0; d < nd; d++) m.push_back({AXES[d], uint32_t(8 * d), h5lite::f64_type()});
    return h5lite::compound_type(uint32_t(8 * nd), m);
}

// prob_domain_t (SayakaWriterHdf5.cpp:28-39): nd low then nd high ints
h5lite::Type box_type(int nd) {
    std::vector<h5lite::Member> m;
    for (int d = 0; d < nd; d++) m.push_back({LO[d], uint32_t(4 * d), h5lite::int32_type()});
    for (int d = 0; d < nd; d++) m.push_back({HI[d], uint32_t(4 * (nd + d)), h5lite::int32_type()});
    return h5lite::compound_type(uint32_t(8 * nd), m);
}

h5lite::Attribute bytes_attr(const std::string& name, const h5lite::Type& t, const void* p) {
    const uint8_t* q = static_cast<const uint8_t*>(p);
    return h5lite::Attribute{name, t, std::vector<uint8_t>(q, q + t.size)};
}

}  // namespace

int WriteChomboPacked(const AmrTree& tree, int ncomp, const double* packed, const char* filename,
                      WriterMeta& meta, std::string* err) {
    auto fail = [&](const std::string& m) {
        if (err) *err = m;
        return -1;
    };
    if (ncomp < 1 || !packed || !filename) return fail("bad argument");
    if (tree.numBlocks < 1) return fail("empty tree");
    const int nd = tree.ndim;
    const int ndata = tree.cellsPerBlock();
    const int num_levels = tree.currentFinestLevel();
    meta.guessName(ncomp);

    // cell size per level: block 0's, halved per level (SayakaWriterHdf5.cpp:368-377)
    std::vector<Vector3d> dx_level(num_levels);
    dx_level[0] = tree.getBlockCellSize(0);
    for (int l = 1; l < num_levels; l++) {
        dx_level[l] = dx_level[l - 1];
        for (int d = 0; d < nd; d++) dx_level[l](d) /= 2.0;
    }
    // problem box: the union of every block's box (SayakaWriterHdf5.cpp:389-405)
    double problo[3] = {0, 0, 0}, probhi[3] = {0, 0, 0};
    for (int b = 0; b < tree.numBlocks; b++) {
        const RealBox& bb = tree[b].boundBox;
        for (int d = 0; d < nd; d++) {
            problo[d] = b == 0 ? bb.vlo(d) : std::min(problo[d], bb.vlo(d));
            probhi[d] = b == 0 ? bb.vhi(d) : std::max(probhi[d], bb.vhi(d));
        }
    }

    const h5lite::Type ti = h5lite::int32_type(), tf = h5lite::f64_type();
    const h5lite::Type trv = realvect_type(nd), tbox = box_type(nd);
    h5lite::Group root;
    root.attrs.push_back(h5lite::scalar_attr("num_components", ti, int32_t(ncomp)));
    root.attrs.push_back(h5lite::scalar_attr("num_levels", ti, int32_t(num_levels)));
    root.attrs.push_back(h5lite::scalar_attr("iteration", ti, int32_t(meta.step)));
    root.attrs.push_back(h5lite::scalar_attr("time", tf, meta.time));
    {
        const Vector3d dh = tree.getBlockCellSize(0);
        const double ar[3] = {1.0, dh.y / dh.x, dh.z / dh.x};
        root.attrs.push_back(bytes_attr("aspect_ratio", trv, ar));
    }
    root.attrs.push_back(h5lite::scalar_attr("data_centering", ti, int32_t(0)));
    root.attrs.push_back(bytes_attr("prob_lo", trv, problo));
    for (int c = 0; c < ncomp; c++) {
        // the first 64 bytes of the name, NUL-padded (the reference hands H5Awrite the
        // string's c_str() for a 64-byte type, SayakaWriterHdf5.cpp:200-216)
        char name[64];
        std::memset(name, 0, sizeof(name));
        std::memcpy(name, meta.name[c].data(), std::min<size_t>(meta.name[c].size(), sizeof(name)));
        root.attrs.push_back(bytes_attr("component_" + std::to_string(c), h5lite::string_type(64), name));
    }
    h5lite::Group glob;
    glob.name = "Chombo_global";
    glob.attrs.push_back(h5lite::scalar_attr("SpaceDim", ti, int32_t(nd)));
    glob.attrs.push_back(h5lite::scalar_attr("testReal", tf, 0.0));
    root.groups.push_back(glob);

    std::vector<std::vector<int32_t>> boxes(num_levels);
    size_t off = 0;
    for (int il = 0; il < num_levels; il++) {
        const std::vector<int>& lv = tree.getTreeLevel(il + 1).blockIds();
        const Vector3d dh = dx_level[il];
        h5lite::Group g;
        g.name = "level_" + std::to_string(il);
        g.attrs.push_back(h5lite::scalar_attr("ref_ratio", ti, int32_t(2)));
        g.attrs.push_back(h5lite::scalar_attr("dt", tf, meta.dt));
        g.attrs.push_back(h5lite::scalar_attr("dx", tf, dh.x));
        int32_t dom[6];
        for (int d = 0; d < nd; d++) {   // SayakaWriterHdf5.cpp:272-281
            dom[d] = 0;
            dom[nd + d] = int32_t((probhi[d] - problo[d] - 0.5 * dh(d)) / dh(d));
        }
        g.attrs.push_back(bytes_attr("prob_domain", tbox, dom));
        std::vector<int32_t>& bx = boxes[il];
        bx.resize((size_t)2 * nd * lv.size());
        for (size_t q = 0; q < lv.size(); q++) {   // SayakaWriterHdf5.cpp:449-466
            const RealBox& bb = tree[lv[q]].boundBox;
            for (int d = 0; d < nd; d++) {
                bx[q * 2 * nd + d] = int32_t((bb.vlo(d) + 0.5 * dh(d) - problo[d]) / dh(d));
                bx[q * 2 * nd + nd + d] = int32_t((bb.vhi(d) - 0.5 * dh(d) - problo[d]) / dh(d));
            }
        }
        g.datasets.push_back(h5lite::Dataset{"boxes", tbox, lv.size(), bx.data()});
        const uint64_t n = (uint64_t)ndata * ncomp * lv.size();
        g.datasets.push_back(h5lite::Dataset{"data:datatype=0", tf, n, packed + off});
        off += n;
        h5lite::Group da;
        da.name = "data_attributes";
        da.attrs.push_back(h5lite::scalar_attr("comps", ti, int32_t(ncomp)));
        g.groups.push_back(da);
        root.groups.push_back(g);
    }
    std::string e;
    if (h5lite::write(filename, root, &e) != 0) return fail(e);
    return 0;
}

int WriteDataHdf5(const TreeData& data, const char* filename, WriterMeta& meta, std::string* err) {
    if (data.faceDir() >= 0) {   // SayakaWriterHdf5.cpp:158-167
        if (err) *err = "only cell data";
        return -1;
    }
    const AmrTree& tree = data.tree();
    const int ncomp = data.numComp(), nfin = tree.currentFinestLevel();
    const int ex = data.validExtent(0), ey = data.validExtent(1), ez = data.validExtent(2);
    std::vector<double> packed;
    packed.reserve((size_t)tree.numBlocks * ncomp * tree.cellsPerBlock());
    for (int l = 1; l <= nfin; l++)
        for (int b : tree.getTreeLevel(l))
            for (int c = 0; c < ncomp; c++)
                for (int k = 0; k < ez; k++)
                    for (int j = 0; j < ey; j++)
                        for (int i = 0; i < ex; i++) packed.push_back(data(b, i, j, k, c));
    return WriteChomboPacked(tree, ncomp, packed.data(), filename, meta, err);
}

}  // namespace amd
}  // namespace sayaka

namespace {
thread_local std::string g_werr;
}

extern "C" {

const char* sy_writer_last_error(void) { return g_werr.c_str(); }

int sy_write_chombo(const sy_tree* t, int ncomp, const double* packed, const char* filename, double time,
                    double dt, int step, const char* const* names) {
    if (!t) {
        g_werr = "null tree";
        return -1;
    }
    sayaka::WriterMeta meta;
    meta.time = time;
    meta.dt = dt;
    meta.step = step;
    if (names)
        for (int c = 0; c < ncomp; c++) meta.name.push_back(names[c] ? names[c] : "");
    std::string err;
    const int rc = sayaka::WriteChomboPacked(*reinterpret_cast<const sayaka::AmrTree*>(t), ncomp, packed,
                                             filename, meta, &err);
    if (rc != 0) g_werr = err;
    return rc;
}

}  // extern "C"
