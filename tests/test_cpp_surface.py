"""The C++ call surface a driver compiles against (include/sayaka/*.h): a small program using
AmrTree, TreeData (Make/Create*, Copy, AddEqual, SetValue, face arrays), BCPatch, FillPatch,
InterpPatch, TreeStateData and WriterMeta, built with g++ and linked to the host library.  It
runs the host-side parts (no GPU call) and checks TreeData::Copy / AddEqual against their
definition (SayakaTreeData.cpp:230-340)."""
from __future__ import annotations

import os
import subprocess

import pytest

from sayakaamr_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r"""
#include <cstdio>
#include "amr_tree.h"
#include "fill_patch.h"
#include "mac_solver.h"
#include "tree_data.h"
#include "writer.h"
using namespace sayaka;
int main() {
    AmrTree t;
    t.ndim = 2; t.nb = 8; t.blockNumGrow = 2; t.minAmrLevel = 2; t.maxAmrLevel = 3; t.maxBlockNum = 500;
    t.init();
    const int ndiv[3] = {2, 2, 1};
    RealBox box; box.vlo = Vector3d::VecMake(0, 0, 0); box.vhi = Vector3d::VecMake(1, 1, 0);
    const int bc[6] = {-100, -100, -100, -100, -100, -100};
    const bool per[6] = {false, false, false, false, false, false};
    t.initUniformRootLevel(ndiv, box, bc, per);
    t.initRefineToMinLevel();
    TreeData* a = TreeData::CreateCellData(t, 2, 1);
    TreeData* b = TreeData::CreateCellData(t, 3, 2);
    for (int blk = 0; blk < t.numBlocks; blk++)
        for (int j = -1; j < 9; j++)
            for (int i = -1; i < 9; i++) (*a)(blk, i, j, 0, 1) = blk * 100 + j * 10 + i;
    b->SetValue(0.5);
    TreeData::Copy(*b, *a, 2, 1, 1, 1);
    TreeData::AddEqual(*b, *a, 2.0, 0, 1, 1, 0);
    int bad = 0;
    for (int blk = 0; blk < t.numBlocks; blk++)
        for (int j = -2; j < 10; j++)
            for (int i = -2; i < 10; i++) {
                const bool in1 = i >= -1 && i < 9 && j >= -1 && j < 9, in0 = i >= 0 && i < 8 && j >= 0 && j < 8;
                const double v = blk * 100 + j * 10 + i;
                if ((*b)(blk, i, j, 0, 2) != (in1 ? v : 0.5)) bad++;
                if ((*b)(blk, i, j, 0, 0) != (in0 ? 0.5 + 2.0 * v : 0.5)) bad++;
                if ((*b)(blk, i, j, 0, 1) != 0.5) bad++;
            }
    std::vector<TreeData*> faces = TreeData::CreateFaceDataPArray(t, 1, 0);
    if (faces.size() != 2 || faces[0]->validExtent(0) != 9 || faces[1]->validExtent(1) != 9) bad++;
    TreeData::ReleaseDataPArray(faces);
    BCPatch bcp;
    bcp.defineByData(*b);
    bcp.boundaryRegister(0).setBCMap(-1, 0, 0, -100, Neumann);
    FillPatch fp(t, *b);
    fp.setBC(bcp);
    fp.setProlongation(0, InterpPatch::PROLONG_CENTERED_LINEAR);
    TreeStateData tsd;
    tsd.new_data = b;
    tsd.fill_new_patch = &fp;
    t.treeStateData.push_back(tsd);
    WriterMeta meta;
    meta.guessName(*b);
    if (meta.name.size() != 3 || meta.name[2] != "comp2" || meta.dt != 1.0) bad++;
    t.treeStateData.clear();
    delete a;
    delete b;
    std::printf("blocks=%d bad=%d\n", t.numBlocks, bad);
    return bad ? 1 : 0;
}
"""


def test_driver_surface_compiles_and_treedata_utilities(tmp_path):
    src = tmp_path / "driver.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "driver"
    lib = N.LIBDIR
    inc = os.path.join(ROOT, "include")
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", f"-I{inc}", f"-I{os.path.join(inc, 'sayaka')}",
                        str(src), f"-L{lib}", "-lsayaka_host", "-lsayaka_dev", f"-Wl,-rpath,{lib}", "-o", str(exe)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "bad=0" in r.stdout
