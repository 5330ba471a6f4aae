"""CPU tests of libmfs.so's host logic against the oracle (no device calls).

Covers the synthetic generators (bit-exact), the mechanical renumbering and
the portrait (bit-exact DOF indexing and sparsity pattern), and that the
library exports every symbol include/mfs.h declares.
"""
import ctypes
import os

import numpy as np
import pytest

import multiscalefemsolver3d_amd as mfs
import oracle as orc
from multiscalefemsolver3d_amd import problems


def test_library_exports_every_declared_symbol():
    names = mfs.exported_symbols()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(mfs.lib, n)]
    assert missing == []


def test_mesh_generator_matches_oracle():
    for n in (1, 3, 7):
        x, t = mfs.mesh_box(n)
        x2, t2 = orc.mesh_box(n)
        assert np.array_equal(x, x2) and np.array_equal(t, t2)
    x, t = mfs.mesh_box(2, 3, 4, origin=(0.5, -1, 2), h=(0.25, 0.5, 0.125))
    x2, t2 = orc.mesh_box(2, 3, 4, origin=(0.5, -1, 2), h=(0.25, 0.5, 0.125))
    assert np.array_equal(x, x2) and np.array_equal(t, t2)


def test_fields_match_oracle_bit_exact():
    x, t = mfs.mesh_box(8)
    a = mfs.field_inclusions(t, x, 200, 0.02, 0.06, 2, 1.0, 100.0)
    b = orc.field_inclusions(t, x, 200, 0.02, 0.06, 2, 1.0, 100.0)
    assert np.array_equal(a, b) and (a == 100.0).any()
    a = mfs.field_lognormal(5000, 1, 3, 17, np.log(1e9), 1.0)
    b = orc.field_lognormal(5000, 1, 3, 17, np.log(1e9), 1.0)
    assert np.array_equal(a, b)
    ln = np.log(a)
    assert abs(ln.mean() - np.log(1e9)) < 0.05 and abs(ln.std() - 1.0) < 0.05
    a, ma = mfs.field_grf(x, t, 64, 0.05, 1, 1e9, 2e11)
    b, mb = orc.field_grf(x, t, 64, 0.05, 1, 1e9, 2e11)
    assert np.array_equal(a, b) and np.array_equal(ma, mb)
    assert 0.2 < ma.mean() < 0.8


def test_face_selections():
    n = 4
    x, _ = mfs.mesh_box(n)
    for f in range(6):
        nodes = mfs.box_face_nodes(n, n, n, f)
        ax, side = f // 2, f & 1
        assert np.array_equal(nodes, np.nonzero(x[:, ax] == side)[0])
        tris = mfs.box_face_tris(n, n, n, f)
        assert len(tris) == 2 * n * n
        assert (x[tris][..., ax] == side).all()
        areas = 0.5 * np.linalg.norm(np.cross(x[tris[:, 1]] - x[tris[:, 0]], x[tris[:, 2]] - x[tris[:, 0]]), axis=1)
        assert areas.sum() == pytest.approx(1.0, rel=1e-14)


def test_renumbering_matches_oracle():
    p = problems.elastic_cube(6, field="grf")
    a = mfs.renumber(len(p["xyz"]), p["tets"], p["material"], p["fixed_nodes"])
    b = orc.renumber(len(p["xyz"]), p["tets"], p["material"], p["fixed_nodes"])
    assert np.array_equal(a, b)
    assert np.array_equal(np.sort(a), np.arange(len(a)))


def expand_block_pattern(rp, col):
    """MakeCopyMatrix_A_into_B order (Math.h:1963-2054) written from its rule."""
    rows = [0]
    cols = []
    for i in range(len(rp) - 1):
        bc = col[rp[i]:rp[i + 1]]
        for a in range(3):
            cols.append((3 * bc[:, None] + np.arange(3)).ravel())
            rows.append(rows[-1] + 3 * len(bc))
    return np.array(rows, dtype=np.int32), np.concatenate(cols).astype(np.int32)


def test_scalar_portrait_bit_exact_vs_oracle():
    p = problems.laplace_inclusions(6)
    rp, col = mfs.portrait(len(p["xyz"]), p["tets"])
    S = orc.scalar_system(p["xyz"], p["tets"], p["sigma"], p["fixed_nodes"], p["fixed_values"])
    assert np.array_equal(rp, S.rowptr) and np.array_equal(col, S.col)


def test_elastic_dof_indexing_and_pattern_bit_exact_vs_oracle():
    p = problems.elastic_cube(5, field="grf")
    nn = len(p["xyz"])
    m = mfs.renumber(nn, p["tets"], p["material"], p["fixed_nodes"])
    rt = m[p["tets"]]
    rp, col = mfs.portrait(nn, rt)
    srp, scol = expand_block_pattern(rp, col)
    S = orc.elastic_system(p["xyz"], p["tets"], p["E"], p["nu"], p["fixed_nodes"], p["fixed_values"], p["tris"],
                           p["traction"], p["material"])
    assert np.array_equal(S.node_map(nn), m)
    assert np.array_equal(srp, S.rowptr) and np.array_equal(scol, S.col)


def test_elastic20_sizes():
    p = problems.elastic_cube(20)
    rp, col = mfs.portrait(len(p["xyz"]), p["tets"])
    assert len(p["xyz"]) == 9261 and len(p["tets"]) == 48000
    assert 9 * len(col) == 1157229
    assert len(p["fixed_nodes"]) == 441 and len(p["tris"]) == 800


def test_gmsh_writer_layout(tmp_path):
    """Lines follow the 'modified' layout SimpleGrid::ReadFromMSH parses
    (Library/Math.h:327-520): header '2 0 8 1', 9-int triangle rows with the
    parent tet in the 5th field, 10-int tet rows with the material in the 4th."""
    x, t = mfs.mesh_box(2)
    tris = mfs.box_face_tris(2, 2, 2, 5)
    path = str(tmp_path / "m.msh")
    mfs.write_gmsh(path, x, t, np.zeros(len(t), np.int32), tris, np.full(len(tris), 3, np.int32))
    lines = open(path).read().split("\n")
    assert lines[1].split() == ["2", "0", "8", "1"]
    assert int(lines[4]) == len(x)
    el = [list(map(int, l.split())) for l in lines[4 + len(x) + 4:] if l.strip() and not l.startswith("$")]
    tri_rows = [r for r in el if len(r) == 9]
    tet_rows = [r for r in el if len(r) == 10]
    assert len(tri_rows) == len(tris) and len(tet_rows) == len(t)
    for r, tri in zip(tri_rows, tris):
        assert r[1] == 2 and r[3] == 3
        parent = t[r[4] - 1]
        assert set((np.array(r[6:9]) - 1).tolist()) <= set(parent.tolist())
    assert all(r[1] == 4 and r[3] == 0 for r in tet_rows)


def test_gmsh_reader_round_trip_and_parent_search(tmp_path):
    """The reader returns what the writer wrote (coordinates to the written
    16 decimals, everything else exactly, parents included); a file in the
    plain layout (header "2.2 0 8", no parent field) gets its parents from
    the adjacency, checked here by brute force."""
    x, t = mfs.mesh_box(3)
    mat = (np.arange(len(t)) % 3).astype(np.int32)
    tris = np.concatenate([mfs.box_face_tris(3, 3, 3, f) for f in (0, 5)])
    tags = np.repeat(np.array([1, 6], np.int32), len(tris) // 2)
    path = str(tmp_path / "rt.msh")
    mfs.write_gmsh(path, x, t, mat, tris, tags)
    x2, t2, m2, tr2, tg2, par = mfs.read_gmsh(path)
    assert np.abs(x2 - x).max() < 1e-15
    assert np.array_equal(t2, t) and np.array_equal(m2, mat)
    assert np.array_equal(tr2, tris) and np.array_equal(tg2, tags)
    for f, p_ in zip(tris, par):
        assert set(f.tolist()) <= set(t[p_].tolist())
    # plain layout: rewrite the header and drop the parent (5th field -> 0)
    lines = open(path).read().split("\n")
    lines[1] = "2.2 0 8"
    out = []
    for l in lines:
        f = l.split()
        if len(f) == 9 and f[1] == "2":
            f[4] = "0"
            l = "\t".join(f)
        out.append(l)
    plain = str(tmp_path / "plain.msh")
    open(plain, "w").write("\n".join(out))
    _, _, _, tr3, _, par3 = mfs.read_gmsh(plain)
    assert np.array_equal(tr3, tris)
    for f, p_ in zip(tris, par3):
        owners = [k for k in range(len(t)) if set(f.tolist()) <= set(t[k].tolist())]
        assert p_ in owners


def test_gmsh_reader_rejects_bad_input(tmp_path):
    bad = tmp_path / "bad.msh"
    bad.write_text("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n2\n1 0 0 0\n5 1 1 1\n$EndNodes\n")
    with pytest.raises(Exception):
        mfs.read_gmsh(str(bad))
    with pytest.raises(Exception):
        mfs.read_gmsh(str(tmp_path / "missing.msh"))
    # negative and absurd $Nodes counts are malformed files, not allocations
    for count in ("-5", "99999999999999", "x"):
        huge = tmp_path / f"count{count}.msh"
        huge.write_text(f"$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n{count}\n1 0 0 0\n$EndNodes\n")
        with pytest.raises(mfs.MfsError, match="malformed"):
            mfs.read_gmsh(str(huge))


def parse_params_like_the_reference(text):
    """MechanicalMain.cpp:859-1201 restated: first line by %d; then lines up
    to '\n' or '#', empty ones skipped; the tag is the line's first
    character; a Dirichlet value counts only when a space follows it, '*'
    frees a component; Neumann rows are value + direction."""
    lines = text.split("\n")
    first = int(lines[0].split()[0])
    rest = [l.split("#")[0] for l in lines[1:]]
    rest = [l for l in rest if l.strip()]
    it = iter(rest)
    out = {"log": first, "threads": int(next(it).split()[0]), "format": int(next(it).split()[0]),
           "mesh": next(it), "residual": float(next(it).split()[0])}
    nd = int(next(it).split()[0])
    out["dirichlet"] = []
    for _ in range(nd):
        line = next(it)
        tag = int(line[0])
        vals, tok, jj = [None, None, None], "", 0
        for ch in line[1:]:
            if ch == "*":
                jj += 1
            elif ch == " ":
                if tok:
                    vals[jj] = float(tok)
                    jj += 1
                    tok = ""
            elif ch != "\t":
                tok += ch
        out["dirichlet"].append((tag, tuple(vals)))
    nn = int(next(it).split()[0])
    out["neumann"] = []
    for _ in range(nn):
        line = next(it)
        f = [float(v) for v in line[1:].split()]
        out["neumann"].append((int(line[0]), f[0], tuple(f[1:4])))
    nm = int(next(it).split()[0])
    out["materials"] = [tuple(float(v) for v in next(it).split()[:2]) for _ in range(nm)]
    out["result"] = next(it)
    return out


def test_solver_params_writer_reads_back_with_the_reference_rules(tmp_path):
    path = str(tmp_path / "param_for_solver_0.txt")
    dirichlet = [(5, (0.0, 0.0, 0.0)), (3, (None, 1.25e-3, None))]
    neumann = [(6, 1e6, (0.0, 0.0, -1.0)), (2, 2.5e5, (0.0, 0.0, 0.0))]
    materials = [(1e9, 0.3), (2e11, 0.25)]
    mfs.write_solver_params(path, "/data/mesh/cube.msh", "/data/out", dirichlet, neumann, materials, threads=16,
                            critical_residual=1e-8)
    got = parse_params_like_the_reference(open(path).read())
    assert got["log"] == 0 and got["threads"] == 16 and got["format"] == 10
    assert got["mesh"] == "/data/mesh/cube.msh" and got["result"] == "/data/out"
    assert got["residual"] == 1e-8
    assert got["dirichlet"] == dirichlet
    assert got["neumann"] == neumann
    assert got["materials"] == materials
    with pytest.raises(Exception):
        mfs.write_solver_params(path, "m.msh", "out", [(12, (0, 0, 0))])


def test_error_reporting_without_device():
    code = mfs.lib.mfs_mesh_box_counts(0, 1, 1, None, None)
    assert code == -1
    assert b"bad argument" in mfs.lib.mfs_last_error()
