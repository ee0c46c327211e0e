"""The reference-shaped C++ host API (include/graph.h, include/pattern.hh)
built as a caller would build it: C++ programs compiled against include/ and
linked with libjgpm.so.

* tests/cpp/api_members.cc drives the Graph and Pattern members with no
  reference-free counterpart elsewhere (COO task list, connectivity of a
  vertex set, degree relabelling, the reference's pattern names); its output
  is compared with numpy restatements here.
* tests/cpp/integration_main.cc is the reference-layout main.cc of the
  integration notes; it is linked here and run on the GPU on a graph file,
  where its two result lines must carry the library's counts.
* The file loader rejects lists that are unsorted, duplicated, self-looped
  or (for the symmetric loaders) asymmetric.
"""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

import jupiter_amd as J

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "tests", "cpp")


def build(tmp_path, name):
    cxx = shutil.which("g++") or shutil.which("c++")
    if cxx is None:
        pytest.skip("no host C++ compiler")
    exe = str(tmp_path / name)
    lib_dir = os.path.dirname(J.LIB_PATH)
    subprocess.check_call([cxx, "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "include"), "-o", exe,
                           os.path.join(CPP, name + ".cc"), "-L" + lib_dir, "-ljgpm", "-Wl,-rpath," + lib_dir])
    return exe


def test_graph_and_pattern_members(tmp_path):
    exe = build(tmp_path, "api_members")
    (tmp_path / "w.txt").write_text("0 1\n1 2\n")
    (tmp_path / "c4.txt").write_text("0 1\n1 2\n2 3\n3 0\n")
    out = subprocess.run([exe, str(tmp_path / "w.txt"), str(tmp_path / "c4.txt")], capture_output=True,
                         text=True, timeout=60, check=True).stdout
    lines = {ln.split()[0]: ln.split()[1:] for ln in out.splitlines() if ln.strip()}
    rp = np.array([0, 2, 4, 7, 9, 10, 10])
    col = np.array([1, 2, 0, 2, 0, 1, 3, 2, 4, 3])
    deg = np.diff(rp)
    assert lines["tasks_before"] == ["0"] and lines["tasks"] == ["10", "10"]
    assert list(map(int, lines["src"])) == np.repeat(np.arange(6), deg).tolist()
    assert list(map(int, lines["dst"])) == col.tolist()
    assert lines["src_ptr_ok"] == ["1"]
    # {0..4} connected; {0,3} not adjacent; {4,5}: 5 isolated; one vertex is connected
    assert lines["connected"] == ["1", "0", "0", "1"]
    assert lines["N2"] == ["3:", "0", "1", "3"]
    # degree relabelling: stable by id among equal degrees, lists re-sorted
    order = sorted(range(6), key=lambda v: (deg[v], v))
    rank = {v: k for k, v in enumerate(order)}
    new_rp = [0]
    new_col = []
    for v in order:
        lst = sorted(rank[int(u)] for u in col[rp[v]:rp[v + 1]])
        new_col += lst
        new_rp.append(new_rp[-1] + len(lst))
    assert list(map(int, lines["sorted_rowptr"])) == new_rp
    assert list(map(int, lines["sorted_col"])) == new_col
    assert "vertex 5: degree = 3 edgelist = [ 2 3 4 ]" in out
    # the reference's names: a 2-edge 3-vertex pattern is a wedge, any
    # 4-edge 4-vertex pattern a "square" (common/pattern.cc:3-27)
    assert lines["pattern"][:2] == ["wedge", "wedge"] and lines["pattern"][2] == "1"
    assert lines["pattern"][3:] == ["triangle", "0", "diamond", "0", "rect", "0", "pent", "0", "house", "0",
                                    "sq4", "0"]
    assert lines["v_list"] == ["0", "1", "2"]
    assert lines["neighbor"] == ["1", "-1"]
    assert "|V|: 3, |E|: 2" in out
    assert lines["pattern2"] == ["square", "1"]


def test_integration_main_links(tmp_path):
    assert os.access(build(tmp_path, "integration_main"), os.X_OK)


def _write_raw(prefix, rp, col):
    rp = np.asarray(rp, dtype=np.int64)
    col = np.asarray(col, dtype=np.int32)
    rp.tofile(prefix + ".vertex.bin")
    col.tofile(prefix + ".edge.bin")
    with open(prefix + ".meta.txt", "w") as f:
        f.write(f"{len(rp) - 1}\n{len(col)}\n4 8 1 2\n{int(np.diff(rp).max())}\n0\n0\n0\n")


@pytest.mark.parametrize("case,rp,col", [
    ("unsorted", [0, 2, 3, 4], [2, 1, 0, 0]),
    ("duplicate", [0, 2, 3, 3], [1, 1, 0]),
    ("self loop", [0, 2, 3, 4], [0, 1, 0, 0]),
    ("out of range", [0, 1, 2], [1, 7]),
])
def test_loader_rejects_malformed_lists(tmp_path, case, rp, col):
    prefix = str(tmp_path / "bad")
    _write_raw(prefix, rp, col)
    with pytest.raises(J.JgpmError) as e:
        J.CSR.read(prefix)
    assert e.value.status == -5, case


def test_loader_keeps_a_dag_file(tmp_path):
    # an oriented graph written verbatim loads as it is (Graph::dump_binary)
    prefix = str(tmp_path / "dag")
    _write_raw(prefix, [0, 2, 3, 3], [1, 2, 2])
    h = J.CSR.read(prefix)
    assert h.col.tolist() == [1, 2, 2]


@pytest.mark.gpu
def test_symmetric_loaders_reject_asymmetric_lists(ctx, tmp_path):
    prefix = str(tmp_path / "asym")
    _write_raw(prefix, [0, 2, 3, 4], [1, 2, 0, 1])  # 0->2 without 2->0
    print("[case] asymmetric file load", flush=True)
    with pytest.raises(J.JgpmError) as e:
        ctx.load(prefix)
    assert e.value.status == -5
    with pytest.raises(J.JgpmError) as e:
        ctx.upload(J.CSR(np.array([0, 2, 3, 4]), np.array([1, 2, 0, 1])))
    assert e.value.status == -1


@pytest.mark.gpu
def test_integration_main_runs(ctx, tmp_path):
    exe = build(tmp_path, "integration_main")
    g = J.rmat(9, 16, seed=1)
    prefix = str(tmp_path / "graph")
    g.write(prefix)
    G = ctx.upload(g)
    try:
        for pat in ("P13", "P1", "P2"):
            print(f"[case] integration main {pat}", flush=True)
            r = subprocess.run([exe, prefix, pat, "2"], capture_output=True, text=True, timeout=120)
            assert r.returncode == 0, r.stderr
            m = re.search(r"gpu count:(\d+)\s+runtime:([0-9.]+)s", r.stdout)
            assert m, r.stdout
            k = re.search(r"GPU:0 Kernel runtime = \S+ sec << total_num_triangles = (\d+)", r.stdout)
            assert k and k.group(1) == m.group(1)
            assert int(m.group(1)) == G.count(pat)[0]
    finally:
        G.free()


@pytest.mark.gpu
def test_induced_ids_reject_partial_ranges(ctx):
    # the induced motifs combine counts credited to different vertices; on K4
    # with roots [0, 1) the partial P2 and P13 counts alone give 0 - 6 * 1
    from graphs import complete_graph
    rp, col, _ = complete_graph(4)
    G = ctx.upload(J.CSR(rp, col))
    try:
        for name in ("I_DIAMOND", "I_C4", "I_TAILTRI", "I_STAR3"):
            print(f"[case] {name} on K4 roots [0,1)", flush=True)
            with pytest.raises(J.JgpmError) as e:
                G.count(name, (0, 1))
            assert e.value.status == -1
            assert G.count(name)[0] == 0
        assert G.count("I_K4")[0] == 1 and G.count("I_K4", (0, 1))[0] == 1
    finally:
        G.free()
