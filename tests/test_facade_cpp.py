"""The C++ facade compiles against the reference's class API and links to the
library: a caller's translation unit written for cvt::PMHuberStereo /
cvt::Image builds with g++; host-memory Image save/load runs without a GPU
(PFM bytes and a CVTRaw round trip)."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SRC = r'''
#include "pmh_stereo.h"
#include <cstdio>
#include <vector>
int main(int argc, char** argv) {
    using namespace cvt;
    Image img(5, 3, IFormat::GRAY_FLOAT);              // host image
    std::vector<float> v(15);
    for (int i = 0; i < 15; ++i) v[i] = 0.5f * i;
    img.upload(v.data());
    img.save(std::string(argv[1]) + "/a.pfm");
    img.save(std::string(argv[1]) + "/a.cvtraw");
    Image back;
    back.load(std::string(argv[1]) + "/a.cvtraw");
    std::vector<float> w(15);
    back.download(w.data());
    for (int i = 0; i < 15; ++i) if (w[i] != v[i]) return 2;
    if (back.width() != 5 || back.height() != 3 || back.format() != IFormat::GRAY_FLOAT) return 3;
    try { img.save(std::string(argv[1]) + "/a.png"); return 4; } catch (const Exception&) {}
    if (argc > 2) {   // the stereo class: only constructed when a GPU is asked for
        PMHuberStereo pmh;
        Image dmap;
        pmh.depthMap(dmap, img, img, 1, 4.0f, 1, 4);
        (void)pmh;
    }
    std::puts("ok");
    return 0;
}
'''


def test_facade_compiles_links_and_saves(tmp_path):
    src = tmp_path / "app.cpp"
    src.write_text(SRC)
    exe = tmp_path / "app"
    lib = os.path.join(ROOT, "cvt_amd")
    subprocess.run(["g++", "-std=c++17", "-O1", str(src), "-I", os.path.join(ROOT, "include"),
                    "-I", os.path.join(ROOT, "cvt_amd", "csrc"), "-L", lib, "-lpmh_amd",
                    "-Wl,-rpath," + lib, "-o", str(exe)], check=True)
    out = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True)
    assert out.returncode == 0, (out.returncode, out.stderr)
    assert out.stdout.strip() == "ok"
    b = open(tmp_path / "a.pfm", "rb").read()
    head = b"Pf\n5 3\n-1.0\n"
    v = (0.5 * np.arange(15, dtype=np.float32)).reshape(3, 5)
    assert b == head + v[::-1].astype("<f4").tobytes()
