"""The C++ layer (include/mv_ncc.hpp) on the GPU: the prebuilt check program
(tests/cpp/ncc_cpp_check, built with the library) runs FindNccModel through
mvkit's object, procedural and image-member forms (which it requires to agree
exactly) and FindNccModels over two models (the second restored from the
first's serialized bytes).  Its matches equal, bit for bit, what the C ABI
returns for the same template and image (mv_find_ncc_model,
mv_find_ncc_models), so the layer adds nothing and drops nothing; the C ABI
itself is held to the oracle by the other GPU tests."""
import math
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROGRAM = os.path.join(ROOT, "tests", "cpp", "ncc_cpp_check")


def test_cpp_find_equals_c_abi(capi, scene_small, tmp_path):
    assert os.path.exists(PROGRAM), "tests/cpp/ncc_cpp_check is built by __graft_entry__.build()"
    s = scene_small
    start, extent, min_score, n = -math.radians(20), math.radians(40), 0.7, 2
    tp, ip = tmp_path / "t.u8", tmp_path / "i.u8"
    tp.write_bytes(np.ascontiguousarray(s.template).tobytes())
    ip.write_bytes(np.ascontiguousarray(s.image).tobytes())
    print(f"C++ API find: template {s.template.shape}, image {s.image.shape}", flush=True)
    sys.stdout.flush()
    r = subprocess.run([PROGRAM, "find", str(tp), str(s.template.shape[1]), str(s.template.shape[0]), str(ip),
                        str(s.image.shape[1]), str(s.image.shape[0]), repr(start), repr(extent), repr(min_score),
                        str(n)], capture_output=True, text=True, timeout=100)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.splitlines()
    assert lines[-1] == "all ok"
    got = np.array([[float(x) for x in l.split()[1:]] for l in lines if l.startswith("find ")]).reshape(-1, 4)
    got_m = np.array([[float(x) for x in l.split()[1:]] for l in lines if l.startswith("models ")]).reshape(-1, 5)
    model = capi.NccModel(s.template, 0, start, extent)
    want = model.find(s.image, start, extent, min_score, n, 0.5, "true", 0)
    assert len(want) == n
    assert np.array_equal(got, want), (got, want)
    want_m = capi.find_models([model, model], s.image, start, extent, min_score, n, 0.5, "true", 0)
    assert np.array_equal(got_m, want_m), (got_m, want_m)
