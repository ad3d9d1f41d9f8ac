"""Host-side checks that need no GPU: the C ABI loads and exports every
declared symbol, the reference file formats parse as the reference reads
them (src/main.cpp:60-139), and the closure entry point fails loudly without
a GPU."""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, load_golden

import graspan_g as gg
from graspan_g import gen

HEADER = os.path.join(ROOT, "include", "graspan_g.h")


def test_library_exports_every_declared_symbol():
    text = open(HEADER).read()
    declared = set(re.findall(r"^\s*(?:const\s+)?[\w\s\*]*?\b(gg_\w+)\s*\(", text, re.M))
    assert declared == set(gg.EXPORTED)
    lib = gg.lib()
    for name in declared:
        assert hasattr(lib, name), name


def test_sample_grammar_as_reference_reads_it():
    c = gg.Context()
    c.load_grammar(os.path.join(GOLDEN, "sample_grammar.txt"))
    assert c.grammar_info() == (12, 13)  # NUM RULES: 12, NUM LABELS: 13
    names = [c.label_name(i) for i in range(1, 14)]
    assert names == ["M", "DV", "d", "-d", "V", "MAM", "AMs", "MAs", "Mq", "MA", "-a", "AM", "a"]
    assert c.rules()[:3].tolist() == [[1, 2, 3], [2, 4, 5], [5, 6, 7]]
    assert c.rules()[4].tolist() == [9, 0, 0]


def test_sample_graph_as_reference_reads_it():
    c = gg.Context()
    c.load_grammar(os.path.join(GOLDEN, "sample_grammar.txt"))
    c.load_graph(os.path.join(GOLDEN, "sample_graph.txt"))
    assert c.graph_info() == (530, 2000)  # NUM VERTICES: 530, INITIAL NUM EDGES: 2000
    d = load_golden(os.path.join(GOLDEN, "sample.json"))
    s, t, l = c.graph_edges()
    np.testing.assert_array_equal(s, d["src"])
    np.testing.assert_array_equal(t, d["dst"])
    np.testing.assert_array_equal(l, d["lab"])


def test_generator_files_round_trip(tmp_path):
    g = gen.alias(2000, seed=1)
    gp, rp = str(tmp_path / "g.txt"), str(tmp_path / "r.txt")
    gen.write_files(g, gp, rp)
    c = gg.Context()
    c.load_grammar(rp)
    assert c.grammar_info() == (9, 9)
    assert [c.label_name(i) for i in range(1, 10)] == ["L", "A", "M", "R", "a", "V", "T", "D", "d"]
    c.load_graph(gp)
    s, t, l = c.graph_edges()
    np.testing.assert_array_equal(s, g.src)
    np.testing.assert_array_equal(t, g.dst)
    np.testing.assert_array_equal(l, g.lab)


BAD_GRAPHS = {
    "unknown_label": "2\t3\n0\t1\te\n1\t2\tz\n",
    "vertex_out_of_range": "1\t3\n0\t3\te\n",
    "short_file": "3\t3\n0\t1\te\n",
    "blank_line": "2\t3\n0\t1\te\n\n1\t2\te\n",
    "bad_header": "2 3\n0\t1\te\n1\t2\te\n",
    "space_separated": "1\t3\n0 1 e\n",
}


@pytest.mark.parametrize("name", sorted(BAD_GRAPHS))
def test_malformed_graph_rejected(tmp_path, name):
    rp = tmp_path / "r.txt"
    rp.write_text("2\nN\te\nN\tN\te\n")
    gp = tmp_path / "g.txt"
    gp.write_text(BAD_GRAPHS[name])
    c = gg.Context()
    c.load_grammar(str(rp))
    with pytest.raises(gg.GGError) as ei:
        c.load_graph(str(gp))
    assert ei.value.status == gg.GG_E_FORMAT


@pytest.mark.parametrize("text", ["x\nN\te\n", "3\nN\te\nN\tN\te\n", "1\nA\tB\tC\tD\n", "1\n\n", ""])
def test_malformed_grammar_rejected(tmp_path, text):
    rp = tmp_path / "r.txt"
    rp.write_text(text)
    c = gg.Context()
    with pytest.raises(gg.GGError) as ei:
        c.load_grammar(str(rp))
    assert ei.value.status == gg.GG_E_FORMAT


def test_comments_skipped_and_unsorted_edges_accepted(tmp_path):
    rp = tmp_path / "r.txt"
    rp.write_text("# c\n2\n# rules\nN\te\n#x\nN\tN\te\n")
    gp = tmp_path / "g.txt"
    gp.write_text("# hdr\n3\t4\n2\t3\te\n# mid\n0\t1\te\n1\t2\te\n")
    c = gg.Context()
    c.load_grammar(str(rp))
    c.load_graph(str(gp))
    s, t, _ = c.graph_edges()
    assert s.tolist() == [2, 0, 1] and t.tolist() == [3, 1, 2]


def test_missing_file_is_io_error():
    c = gg.Context()
    with pytest.raises(gg.GGError) as ei:
        c.load_grammar("/nonexistent/grammar.txt")
    assert ei.value.status == gg.GG_E_IO


def test_set_coo_validates():
    c = gg.Context()
    c.set_grammar(np.array([[1, 2, 0], [1, 1, 2]], np.int32), 2)
    with pytest.raises(gg.GGError):
        c.set_graph(3, np.array([0], np.uint32), np.array([3], np.uint32), np.array([2], np.uint8))
    with pytest.raises(gg.GGError):
        c.set_graph(3, np.array([0], np.uint32), np.array([1], np.uint32), np.array([3], np.uint8))
    with pytest.raises(gg.GGError):
        c.set_grammar(np.array([[1, 0, 2]], np.int32), 2)  # rhs2 without rhs1


def test_run_without_gpu_fails_loudly():
    """No silent CPU fallback: without a GPU the closure is an error."""
    import ctypes
    if _gpu_present():
        pytest.skip("a GPU is present")
    c = gg.Context()
    c.load_grammar(os.path.join(GOLDEN, "sample_grammar.txt"))
    c.load_graph(os.path.join(GOLDEN, "sample_graph.txt"))
    with pytest.raises(gg.GGError) as ei:
        c.run()
    assert ei.value.status == gg.GG_E_DEVICE


def test_row_id_limit_refused_before_device_work():
    """num_labels * num_vertices >= 2^32 cannot be addressed by 32-bit row
    ids: the in-core, partitioned and out-of-core entry points all refuse it
    with GG_E_UNSUPPORTED before they look for a device (so this holds with
    or without a GPU)."""
    g = gen.alias(50, seed=1)
    c = gg.Context()
    c.set_grammar(g.rules, g.num_labels)
    n = (1 << 32) // g.num_labels + 1
    c.set_graph(n, np.array([0], np.uint32), np.array([1], np.uint32), np.array([2], np.uint8))
    for call in (lambda: c.run(0), lambda: c.run(1), lambda: c.run_out_of_core(4)):
        with pytest.raises(gg.GGError) as ei:
            call()
        assert ei.value.status == gg.GG_E_UNSUPPORTED
        assert "2^32" in str(ei.value)


def test_comp_cli_without_gpu_exits_nonzero():
    if _gpu_present():
        pytest.skip("a GPU is present")
    p = subprocess.run([gg.COMP_PATH, os.path.join(GOLDEN, "sample_graph.txt"),
                        os.path.join(GOLDEN, "sample_grammar.txt"), "0"], capture_output=True, text=True)
    assert p.returncode != 0
    assert "NUM COMPUTE-CAPABLE DEVICES: 0" in p.stdout


def _gpu_present():
    return os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK | os.W_OK) and \
        len(glob_dri()) > 0


def glob_dri():
    import glob
    return glob.glob("/dev/dri/renderD*")


def test_triple_hash_definition():
    # splitmix64 finaliser of a linear mix, as stated in the header
    def ref(s, d, l):
        M = (1 << 64) - 1
        x = ((s * 0x9E3779B97F4A7C15) & M) ^ ((d * 0xC2B2AE3D27D4EB4F) & M) ^ ((l * 0x165667B19E3779F9) & M)
        x ^= x >> 30
        x = (x * 0xBF58476D1CE4E5B9) & M
        x ^= x >> 27
        x = (x * 0x94D049BB133111EB) & M
        x ^= x >> 31
        return x
    for t in [(0, 0, 1), (1, 2, 3), (4294967295, 123456, 255)]:
        assert gg.triple_hash(*t) == ref(*t)


@pytest.mark.parametrize("labels", [1, 3])
def test_generator_dedup_wide_vertex_ids(labels):
    """Vertex ids of 2^24 and above keep their value through the
    generators' dedup-and-sort (every id width, one label or several)."""
    rng = np.random.default_rng(7)
    src = rng.integers(0, 1 << 32, 50_000, dtype=np.uint64).astype(np.uint32)
    dst = rng.integers(0, 1 << 32, 50_000, dtype=np.uint64).astype(np.uint32)
    lab = rng.integers(1, labels + 1, 50_000).astype(np.uint8)
    src, dst, lab = (np.concatenate([x, x[:3000]]) for x in (src, dst, lab))
    s, d, l = gen._dedup_sorted(src, dst, lab)
    want = sorted(set(zip(src.tolist(), lab.tolist(), dst.tolist())))
    assert list(zip(s.tolist(), l.tolist(), d.tolist())) == want
    g = gen.dataflow_blocks(17_000_000, seed=3)
    assert int(g.src.max()) >= (1 << 24) and int(g.dst.max()) < g.n
