"""The drop-in command line (jupiter_gpm <graph> <pattern> [rounds]), run as
a child process: its two result lines carry the library's counts, for
generated graphs, graphs in the reference's file format, several rounds,
pattern files isomorphic to a catalogue pattern (the catalogue kernel) and
pattern files outside it (a compiled plan)."""
import os
import re
import subprocess

import pytest

import jupiter_amd as J

pytestmark = pytest.mark.gpu
CLI = os.path.join(os.path.dirname(J.LIB_PATH), "jupiter_gpm")


def run(*args):
    print(f"[case] jupiter_gpm {' '.join(args)}", flush=True)
    r = subprocess.run([CLI, *args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    m = re.search(r"gpu count:(\d+)\s+runtime:([0-9.]+)s", r.stdout)
    assert m, r.stdout
    assert re.search(r"GPU:0 Kernel runtime = \S+ sec << total_num_triangles = (\d+)", r.stdout)
    return int(m.group(1))


@pytest.fixture(scope="module")
def rmat9(ctx):
    g = J.rmat(9, 16, seed=1)
    G = ctx.upload(g)
    yield g, G
    G.free()


def test_generated_graph_and_rounds(rmat9):
    _, G = rmat9
    assert run("rmat:9:1", "P13") == G.count("P13")[0]
    assert run("rmat:9:1", "TC", "4") == G.count("TC")[0]
    assert run("rmat:9:1", "P2", "3") == G.count("P2")[0]


def test_reference_file_format(rmat9, tmp_path):
    g, G = rmat9
    prefix = str(tmp_path / "graph")
    g.write(prefix)
    assert run(prefix, "P1") == G.count("P1")[0]
    assert run(prefix, "P14") == G.count("P14")[0]


def test_pattern_files(rmat9, tmp_path):
    _, G = rmat9
    c4 = tmp_path / "c4.txt"      # a 4-cycle with relabelled vertices: the P1 kernel
    c4.write_text("7 3\n3 9\n9 5\n5 7\n")
    assert run("rmat:9:1", str(c4)) == G.count("P1")[0]
    c5 = tmp_path / "c5.txt"      # a 5-cycle, not in the catalogue: a compiled plan
    c5.write_text("0 1\n1 2\n2 3\n3 4\n4 0\n")
    assert run("rmat:9:1", str(c5)) == G.count_plan(J.Plan(5, [(0, 1), (1, 2), (2, 3), (3, 4), (4, 0)]))[0]
