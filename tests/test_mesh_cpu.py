"""The gmsh 2.2 reader (``pbte_gmsh_read``) against a line-by-line parse of
the same files written here (nodes, element types, physical ids, nodes of
each element, physical names), and its refusals: another format version,
node ids out of order, another element type, fewer than two tags, a short
section, no nodes (``mesh/gmsh_reader.py:45-122``)."""
import numpy as np
import pytest

from conftest import ROOT


def _parse(path):
    lines = [l.strip() for l in open(path).read().splitlines()]
    sec = {}
    k = 0
    while k < len(lines):
        if lines[k].startswith("$") and not lines[k].startswith("$End"):
            name = lines[k][1:]
            j = lines.index(f"$End{name}", k)
            sec.setdefault(name, lines[k + 1:j])
            k = j
        k += 1
    nn = int(sec["Nodes"][0])
    xy = np.array([[float(v) for v in r.split()[1:3]] for r in sec["Nodes"][1:1 + nn]])
    kinds, phys, conn = [], [], []
    for r in sec["Elements"][1:1 + int(sec["Elements"][0])]:
        t = [int(v) for v in r.split()]
        w = 2 if t[1] == 1 else 3
        kinds.append(t[1])
        phys.append(t[3])
        conn.append([v - 1 for v in t[3 + t[2]:3 + t[2] + w]] + [-1] * (3 - w))
    names = {}
    for r in sec.get("PhysicalNames", ["0"])[1:]:
        d, tag, lab = r.split(None, 2)
        names[(int(d), int(tag))] = lab.strip().strip('"')
    return xy, np.array(kinds), np.array(phys), np.array(conn), names


@pytest.mark.parametrize("name", ["A1_Nx6_Ny6", "C1_Nx11_Ny11", "J_Nx6_jit25_s0"])
def test_reader_matches_line_parse(name):
    from pbte_amd.mesh import read_gmsh22
    path = ROOT / "tests/golden/meshes" / f"{name}.msh"
    g = read_gmsh22(path)
    xy, kinds, phys, conn, names = _parse(path)
    np.testing.assert_array_equal(g.nodes, xy)
    np.testing.assert_array_equal(g.elem_type, kinds)
    np.testing.assert_array_equal(g.elem_phyid, phys)
    np.testing.assert_array_equal(g.elem_nodes, conn)
    assert g.physical_names == names


HEAD = "$MeshFormat\n2.2 0 8\n$EndMeshFormat\n"


@pytest.mark.parametrize("text,msg", [
    ("$MeshFormat\n4.1 0 8\n$EndMeshFormat\n$Nodes\n1\n1 0 0 0\n$EndNodes\n", "gmsh format 4.1"),
    ("$MeshFormat\n2.2 1 8\n$EndMeshFormat\n$Nodes\n1\n1 0 0 0\n$EndNodes\n", "binary gmsh file"),
    (HEAD + "$Nodes\n2\n1 0 0 0\n3 1 0 0\n$EndNodes\n", "node ids do not run 1..2"),
    (HEAD + "$Nodes\n1\n1 0 0 0\n$EndNodes\n$Elements\n1\n1 3 2 1 1 1 1 1 1\n$EndElements\n", "is of type 3"),
    (HEAD + "$Nodes\n1\n1 0 0 0\n$EndNodes\n$Elements\n1\n1 2 1 1 1 1 1\n$EndElements\n", "carries 1 tags"),
    (HEAD + "$Nodes\n3\n1 0 0 0\n$EndNodes\n", "announces 3 entries and holds 1"),
    (HEAD, "no $Nodes section"),
])
def test_reader_refusals(tmp_path, text, msg):
    from pbte_amd.mesh import read_gmsh22
    path = tmp_path / "bad.msh"
    path.write_text(text)
    with pytest.raises(ValueError, match=msg.replace("$", r"\$").replace(".", r"\.")):
        read_gmsh22(path)
