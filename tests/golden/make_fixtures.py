#!/usr/bin/env python3
"""Extract the reference's data fixtures for this path into small JSON files.

Reads Feel++'s own fixture meshes (MSH 2.2 text) and writes their node
coordinates and elements, unchanged, as JSON under tests/golden/:
  data/operatorlagrangep1/simplex-3d-lagrange-order2.msh           P2 tet node set, local dof order
  data/operatorlagrangep1/simplex-3d-lagrange-order3-warpblend.msh P3 warp-blend tet nodes (17 digits)
  data/operatorlagrangep1/simplex-3d-lagrange-order3-equispaced.msh
  data/operatorlagrangep1/simplex-2d-lagrange-order2.msh
  data/operatorlagrangep1/simplex-2d-lagrange-order4-{equispaced,fekete}.msh  P4 triangle node sets
  testsuite/feeldiscr/eight_elements_mesh.msh                      8-triangle mesh
and copies a few of Feel++'s own small meshes byte for byte under
tests/golden/msh/, for reading through the product's MSH 2.2 reader:
  testsuite/feeldiscr/eight_elements_mesh.msh          triangles + points + lines
  data/geo/one-elt-meshes-3d/one-elt-real-rotx.msh     one tetrahedron, $PhysicalNames, permuted vertices
  benchmarks/ATTIC/hdg/tetra.msh                       one tetrahedron with its faces
Usage: python tests/golden/make_fixtures.py /root/reference
"""
import json
import os
import shutil
import sys


def read_msh22(path):
    lines = open(path).read().split("\n")
    it = iter(lines)
    nodes, elems = [], []
    for ln in it:
        if ln.strip() == "$Nodes":
            n = int(next(it))
            for _ in range(n):
                t = next(it).split()
                nodes.append([int(t[0])] + [float(v) for v in t[1:4]])
        if ln.strip() == "$Elements":
            n = int(next(it))
            for _ in range(n):
                t = [int(v) for v in next(it).split()]
                num, typ, ntags = t[0], t[1], t[2]
                elems.append({"num": num, "type": typ, "tags": t[3:3 + ntags], "nodes": t[3 + ntags:]})
    return {"nodes": nodes, "elements": elems}


FILES = {
    "nodes_tet_p2.json": "data/operatorlagrangep1/simplex-3d-lagrange-order2.msh",
    "nodes_tet_p3_warpblend.json": "data/operatorlagrangep1/simplex-3d-lagrange-order3-warpblend.msh",
    "nodes_tet_p3_equispaced.json": "data/operatorlagrangep1/simplex-3d-lagrange-order3-equispaced.msh",
    "nodes_tri_p2.json": "data/operatorlagrangep1/simplex-2d-lagrange-order2.msh",
    "eight_elements_mesh.json": "testsuite/feeldiscr/eight_elements_mesh.msh",
    "nodes_tri_p4_equispaced.json": "data/operatorlagrangep1/simplex-2d-lagrange-order4-equispaced.msh",
    "nodes_tri_p4_fekete.json": "data/operatorlagrangep1/simplex-2d-lagrange-order4-fekete.msh",
}

VERBATIM = {
    "eight_elements_mesh.msh": "testsuite/feeldiscr/eight_elements_mesh.msh",
    "one-elt-real-rotx.msh": "data/geo/one-elt-meshes-3d/one-elt-real-rotx.msh",
    "tetra.msh": "benchmarks/ATTIC/hdg/tetra.msh",
}

if __name__ == "__main__":
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    here = os.path.dirname(os.path.abspath(__file__))
    for out, src in FILES.items():
        d = read_msh22(os.path.join(ref, src))
        d["source"] = src
        with open(os.path.join(here, out), "w") as f:
            json.dump(d, f, indent=1)
        print(out, len(d["nodes"]), "nodes", len(d["elements"]), "elements")
    os.makedirs(os.path.join(here, "msh"), exist_ok=True)
    for out, src in VERBATIM.items():
        shutil.copyfile(os.path.join(ref, src), os.path.join(here, "msh", out))
        print("msh/" + out)
