"""Writes quad_q2_6x5.uhm: the leaf elements of the 6 x 5 Q2 quad mesh of
meshes.qp_spec in the reference's ".uhm" text format
(UHM/src/mesh/mesh.cxx:511-579: "UHM", n_nodes, one "id phy ndof p kind"
line per node, n_elements, per element its node count then one "id phy" line
per node), with three nodes that no element holds (ids 9000 and 9001 among
the others, (7, 1) at the end), which the importer drops."""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from meshes import qp_spec  # noqa: E402


def main():
    spec = qp_spec((6, 5), 2)
    lines = ["UHM"]
    nodes = []
    for k, (nid, nd) in enumerate(spec.nodes):
        nodes.append((nid, 0, nd, nd, 0))
        if k == 10:
            nodes += [(9000, 0, 2, 2, 0), (9001, 0, 1, 1, 0)]
    nodes.append((7, 1, 3, 3, 0))
    lines.append(str(len(nodes)))
    lines += ["%d %d %d %d %d" % n for n in nodes]
    lines.append(str(len(spec.elems)))
    for nl in spec.elems:
        lines.append(str(len(nl)))
        lines += ["%d 0" % nid for nid in nl]
    with open(os.path.join(HERE, "quad_q2_6x5.uhm"), "w") as f:
        f.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    main()
