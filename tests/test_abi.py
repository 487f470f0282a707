"""CPU tests of the C-ABI library: exports, lifecycle semantics (reference
src/dplrini.c, src/dplrresizelr.c), host helpers, the synthetic generator,
and loud failure of the GPU entry points when no device is present."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import dplr_amd as D
from dplr_amd.tree import PlrTree

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(dplr\w*)\s*\(", text))


def test_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", D.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    for header in ("dplr.h", "dplrx.h", "dplr_internal.h"):
        missing = declared(header) - exported
        assert not missing, f"{header}: not exported: {sorted(missing)}"


def test_internal_header_forwards(tmp_path):
    """Code written against the reference's private header builds and links
    unchanged: dplr_internal.h brings the node types and the private
    lifecycle routines (reference include/dplr_internal.h:99-115)."""
    src = tmp_path / "uses_internal.c"
    src.write_text(
        "#include <dplr_internal.h>\n"
        "int main(void) {\n"
        "    DPlrMtx a, b, c;\n"
        "    dplrinid(8, &a);\n"
        "    dplrinilr(8, &b, 2);\n"
        "    dplrresizelr(&b, 5);\n"
        "    dplrinih(8, &c);\n"
        "    int ok = a.type == PLR_DMTX && b.type == PLR_LRMTX && b.dlr.capacity >= 5 &&\n"
        "             c.type == PLR_HMTX;\n"
        "    dplrclr(&a); dplrclr(&b); dplrclr(&c);\n"
        "    return ok ? 0 : 1;\n"
        "}\n")
    exe = tmp_path / "uses_internal"
    libdir = os.path.dirname(D.LIB_PATH)
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                    "-L", libdir, "-ldplr_hip", f"-Wl,-rpath,{libdir}"], check=True)
    assert subprocess.run([str(exe)]).returncode == 0


def test_node_layout():
    from dplr_amd.abi import DPlrMtx
    assert ctypes.sizeof(DPlrMtx) == 40
    assert DPlrMtx.type.offset == 0 and DPlrMtx.size.offset == 4
    lr = D.lib.dplrinilr  # noqa: F841  (layout is what the C side writes)
    t = PlrTree()
    D.lib.dplrinilr(64, t.ref, 3)
    raw = ctypes.string_at(ctypes.addressof(t.root), 40)
    assert int.from_bytes(raw[0:4], "little") == 3          # PLR_LRMTX
    assert int.from_bytes(raw[4:8], "little") == 64         # size
    assert int.from_bytes(raw[8:12], "little") == 0         # rank
    assert int.from_bytes(raw[12:16], "little") == 3        # capacity


def test_lifecycle_matches_reference_semantics(orc):
    t = PlrTree()
    D.lib.dplrinilr(10, t.ref, 2)
    lr = t.root.dlr
    assert lr.rank == 0 and lr.capacity == 2
    assert np.all(np.isnan(np.ctypeslib.as_array(lr.u, shape=(20,))))
    lr.u[0] = 1.5
    lr.sig[0] = 2.0
    lr.rank = 1
    # equal request still doubles (reference src/dplrresizelr.c:42,54)
    D.lib.dplrresizelr(t.ref, 2)
    assert t.root.dlr.capacity == 4
    D.lib.dplrresizelr(t.ref, 9)
    assert t.root.dlr.capacity == 9
    D.lib.dplrresizelr(t.ref, 3)
    assert t.root.dlr.capacity == 9
    assert t.root.dlr.u[0] == 1.5 and t.root.dlr.sig[0] == 2.0
    assert np.isnan(t.root.dlr.u[10])
    D.lib.dplrclr(t.ref)
    assert t.root.type == 0
    D.lib.dplrclr(t.ref)  # clearing twice is harmless
    D.lib.dplrinih(8, t.ref)
    assert t.root.type == 1 and t.root.dh.nw.contents.type == 0
    D.lib.dplrinih(7, t.ref)  # odd order: diagnostic, node untouched
    assert t.root.size == 8


def test_dense_expansion_and_queries(orc, ref):
    rng = np.random.default_rng(0)
    A = PlrTree.from_blocks(256, 3, lambda d, n, w: (d + 2 * n + (w == "sw")) % 5, rng)
    assert np.array_equal(A.dense(), A.dense(api=orc))
    r1 = (ctypes.c_int * 3)(-1, -1, -1)
    r2 = (ctypes.c_int * 3)(-1, -1, -1)
    D.lib.dplrrnk(256, A.ref, r1)
    ref["dplrrnk"](256, A.ref, r2)
    assert list(r1) == list(r2)
    assert D.lib.dplrnz(256, A.ref) == ref["dplrnz"](256, A.ref)


def test_generator_is_deterministic_and_orthonormal():
    A1 = PlrTree.random(1024, 4, 16, seed=5)
    A2 = PlrTree.random(1024, 4, 16, seed=5)
    A3 = PlrTree.random(1024, 4, 16, seed=6)
    d1, d2, d3 = A1.dense(), A2.dense(), A3.dense()
    assert np.array_equal(d1, d2) and not np.array_equal(d1, d3)
    for (d, n, w, node) in A1.blocks():
        u, s, v = A1.lr_arrays(node)
        assert u.shape[1] == 16
        assert np.abs(u.T @ u - np.eye(16)).max() < 1e-13
        assert np.abs(v.T @ v - np.eye(16)).max() < 1e-13
        assert np.allclose(s, 10.0 ** (-np.arange(16) / 4.0), rtol=1e-15, atol=0)
    x1 = np.empty(1000)
    x2 = np.empty(1000)
    D.lib.dplrx_gen_normal(1000, 3, 1, x1.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    D.lib.dplrx_gen_normal(1000, 3, 1, x2.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert np.array_equal(x1, x2) and abs(x1.mean()) < 0.15 and abs(x1.std() - 1) < 0.1


def test_nonuniform_tree_rejected(capfd):
    t = PlrTree()
    D.lib.dplrinih(8, t.ref)
    D.lib.dplrinid(4, t.root.dh.nw)
    D.lib.dplrinih(4, t.root.dh.se)
    for c in ("nw", "se"):
        D.lib.dplrinid(2, getattr(t.root.dh.se.contents.dh, c))
    for node, h in ((t.root.dh.ne, 4), (t.root.dh.sw, 4), (t.root.dh.se.contents.dh.ne, 2),
                    (t.root.dh.se.contents.dh.sw, 2)):
        D.lib.dplrinilr(h, node, 0)
    t.m = 8
    y = np.full(8, 7.0)
    out = D.gepmv("N", t, np.ones(8), 1.0, 0.0, y)
    assert np.array_equal(out, y)
    assert "uniform" in capfd.readouterr().err


@pytest.mark.skipif(__import__("torch").cuda.is_available(), reason="checks the no-GPU path")
def test_gpu_entry_fails_loudly_without_device(capfd):
    A = PlrTree.random(256, 2, 4, seed=1)
    y = np.full(256, 3.0)
    out = D.gepmv("N", A, np.ones(256), 1.0, 0.0, y)
    assert np.array_equal(out, y)          # untouched: no silent CPU fallback
    assert "failed" in capfd.readouterr().err


def test_empty_and_copy_match_reference(ref):
    """dplrempty (src/dplrempty.c:18) and dplrcpy (src/dplrcpy.c:19): the same
    node structure as the reference's, copies equal to the source and
    independent of it, and a copy into a tree of another shape reshapes it."""
    from dplr_amd.tree import copy, empty, nodes
    rng = np.random.default_rng(3)
    A = PlrTree.from_blocks(512, 3, lambda d, n, w: (3 * d + n + (w == "ne")) % 9, rng)
    E1, E2 = empty(A), empty(A, api=ref)
    assert nodes(E1) == nodes(E2)
    assert all(r in (-1, 0) for (_, _, r, _) in nodes(E1))
    B1, B2 = copy(A), copy(A, api=ref)
    assert [n[:3] for n in nodes(B1)] == [n[:3] for n in nodes(B2)] == [n[:3] for n in nodes(A)]
    dA = A.dense()
    assert np.array_equal(B1.dense(), dA) and np.array_equal(B2.dense(), dA)
    # independent storage: changing the copy leaves A alone
    node = next(p for (_, _, _, p) in B1.blocks() if p.contents.dlr.rank > 0)
    node.contents.dlr.sig[0] += 1.0
    assert np.array_equal(A.dense(), dA) and not np.array_equal(B1.dense(), dA)
    # a copy into a tree of another structure reshapes it
    C = PlrTree.random(256, 2, 4, seed=9)
    C.m = 512
    copy(A, C)
    assert [n[:3] for n in nodes(C)] == [n[:3] for n in nodes(A)]
    assert np.array_equal(C.dense(), dA)
    for t in (A, E1, E2, B1, B2, C):
        t.free()


def test_plr_file_round_trip(tmp_path, capfd):
    """dplrx_save / dplrx_load: a uniform random tree, a tree of uneven ranks
    (some blocks of rank 0) and a dense one come back node for node and bit
    for bit; truncated, padded and corrupted files are refused."""
    from dplr_amd.tree import nodes
    rng = np.random.default_rng(4)
    trees = [PlrTree.random(1024, 4, 12, seed=3),
             PlrTree.from_blocks(512, 3, lambda d, n, w: (2 * d + n + (w == "sw")) % 6, rng)]
    dense = PlrTree(64)
    D.lib.dplrinid(64, dense.ref)
    vals = rng.standard_normal(64 * 64)
    ctypes.memmove(dense.root.dd.d, vals.ctypes.data, vals.nbytes)
    trees.append(dense)
    for k, t in enumerate(trees):
        path = tmp_path / f"t{k}.plr"
        t.save(path)
        u = PlrTree.load(path)
        assert u.m == t.m
        assert [n[:3] for n in nodes(u)] == [n[:3] for n in nodes(t)]
        assert np.array_equal(u.dense(), t.dense())
        for (_, _, _, p1), (_, _, _, p2) in zip(t.blocks(), u.blocks()):
            for a, b in zip(t.lr_arrays(p1), u.lr_arrays(p2)):
                assert np.array_equal(a, b)
        raw = path.read_bytes()
        for bad in (raw[:-8], raw + b"\0" * 8, raw[:4] + b"XXXX" + raw[8:],
                    raw[:24] + (12345).to_bytes(4, "little") + raw[28:]):
            bp = tmp_path / "bad.plr"
            bp.write_bytes(bad)
            v = PlrTree(0)
            assert D.lib.dplrx_load(str(bp).encode(), v.ref, None) < 0
            assert v.root.type == 0        # left empty
        u.free()
        t.free()
    assert "not a readable PLR file" in capfd.readouterr().err
