// writer.h -- Chombo-format HDF5 output of cell data (WriterMeta, WriteDataHdf5;
// SayakaWriter.h:13-53, SayakaWriterHdf5.cpp:17-530).  The file is written by the build's own
// HDF5 writer (h5lite), no HDF5 library needed.
#pragma once

#include <string>
#include <vector>

#include "tree_data.h"

namespace sayaka {
inline namespace amd {

// WriterMeta (SayakaWriter.h:13-48)
struct WriterMeta {
    double time;
    double dt;
    int step;
    std::vector<std::string> name;

    WriterMeta() { reset(); }
    void guessName(const TreeData& data) { guessName(data.numComp()); }
    void guessName(int ncomp) {   // "comp<i>" for every unnamed component
        name.resize(ncomp);
        for (int c = 0; c < ncomp; c++)
            if (name[c].empty()) name[c] = "comp" + std::to_string(c);
    }
    void reset() {
        time = 0;
        dt = 1.0;
        step = 0;
        name.clear();
    }
};

// WriteDataHdf5 (SayakaWriterHdf5.cpp:351-530): every block of tree levels 1..finest, valid
// cells of every component.  Cell data only (the reference exits for other locations; here
// -1).  Returns 0, or -1 with *err set when err is given.
int WriteDataHdf5(const TreeData& data, const char* filename, WriterMeta& meta, std::string* err = nullptr);

// The same file from data already in the file's order: for every level 1..finest, every block
// of the level in ascending id, every component, the block's valid cells (i fastest) -- the
// 'data:datatype=0' datasets of all levels back to back (what the device pack writes).
int WriteChomboPacked(const AmrTree& tree, int ncomp, const double* packed, const char* filename,
                      WriterMeta& meta, std::string* err = nullptr);

}  // namespace amd
}  // namespace sayaka
