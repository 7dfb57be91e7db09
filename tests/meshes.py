"""Mesh descriptions shared by the tests: one spec drives both the product's
mesh API (uhm_amd.Mesh) and the CPU oracle, plus an independent formulation
of the lock rules used to pin the oracle's analysis."""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


class MeshSpec:
    """nodes: list of (id, n_dof) in creation order; elems: node id lists
    (user order) of the leaves; parent/children over all elements (leaves
    first, parents after, in creation order)."""

    def __init__(self, nodes, elems, parent, children):
        self.nodes = list(nodes)
        self.elems = [list(e) for e in elems]
        self.parent = list(parent)
        self.children = [list(c) for c in children]
        self.seq = {nid: k for k, (nid, _) in enumerate(self.nodes)}
        self.ndof = {nid: d for nid, d in self.nodes}

    @property
    def n_elem(self):
        return len(self.parent)

    def build_product(self, U):
        m = U.Mesh()
        nh = {}
        for nid, d in self.nodes:
            nh[nid] = m.add_node(nid, d)
        E = [m.add_element() for _ in range(self.n_elem)]
        for e, nl in enumerate(self.elems):
            for nid in nl:
                E[e].add_node(nh[nid])
        for e in range(self.n_elem):
            if self.parent[e] >= 0:
                E[e].set_parent(E[self.parent[e]])
            for c in self.children[e]:
                E[e].add_child(E[c])
        return m, E

    def oracle_arrays(self):
        ndof = np.array([d for _, d in self.nodes], np.int32)
        kind = np.zeros(len(self.nodes), np.int32)
        eptr, enode = [0], []
        for e in range(self.n_elem):
            if e < len(self.elems):
                enode += [self.seq[n] for n in self.elems[e]]
            eptr.append(len(enode))
        cptr, cidx = [0], []
        for e in range(self.n_elem):
            cidx += self.children[e]
            cptr.append(len(cidx))
        return ndof, kind, np.array(eptr, np.int32), np.array(enode, np.int32), np.array(self.parent, np.int32), \
            np.array(cptr, np.int32), np.array(cidx, np.int32)

    def oracle(self):
        from oracle.oracle import OracleMesh
        return OracleMesh(*self.oracle_arrays())


def nd_tree(dims):
    """Geometric nested dissection of the element box: longest axis (first
    on ties) split at lo + len/2, parents created in post order after all
    cells.  Returns (parent, children) over cells + parents."""
    d = len(dims)
    ncell = int(np.prod(dims))
    parent = [-1] * ncell
    children = [[] for _ in range(ncell)]

    def cell(lo):
        idx, stride = 0, 1
        for a in range(d):
            idx += lo[a] * stride
            stride *= dims[a]
        return idx

    def rec(lo, hi):
        lens = [hi[a] - lo[a] for a in range(d)]
        best = max(lens)
        if best == 1:
            return cell(lo)
        axis = lens.index(best)
        mid = lo[axis] + lens[axis] // 2
        hiL = list(hi)
        hiL[axis] = mid
        loR = list(lo)
        loR[axis] = mid
        a = rec(lo, hiL)
        b = rec(loR, hi)
        p = len(parent)
        parent.append(-1)
        children.append([a, b])
        parent[a] = p
        parent[b] = p
        return p

    if ncell > 1:
        rec([0] * d, list(dims))
    return parent, children


def qp_spec(dims, p):
    """Uniform Q_p mesh: node id = grid index (x fastest), 1 dof per node,
    element node list in tensor order (x fastest)."""
    d = len(dims)
    X = [dims[a] * p + 1 for a in range(d)]
    nn = int(np.prod(X))
    nodes = [(i, 1) for i in range(nn)]
    elems = []
    if d == 2:
        for ey in range(dims[1]):
            for ex in range(dims[0]):
                elems.append([(ex * p + a) + X[0] * (ey * p + b) for b in range(p + 1) for a in range(p + 1)])
    else:
        for ez in range(dims[2]):
            for ey in range(dims[1]):
                for ex in range(dims[0]):
                    elems.append([(ex * p + a) + X[0] * ((ey * p + b) + X[1] * (ez * p + c))
                                  for c in range(p + 1) for b in range(p + 1) for a in range(p + 1)])
    parent, children = nd_tree(dims)
    return MeshSpec(nodes, elems, parent, children)


def ex1_spec():
    g = json.load(open(os.path.join(GOLDEN, "ex1_maps.json")))
    nodes = []
    for el in g["elements"]:
        for nid in el:
            if nid not in [n for n, _ in nodes]:
                nodes.append((nid, g["n_dof"][str(nid)]))
    return MeshSpec(nodes, g["elements"], [2, 2, -1], [[], [], [0, 1]]), g


def spec_from_connectivity(c):
    """MeshSpec of a mesh built by the product's synthetic generator."""
    ids = [int(v) for v in c["node_id"]]
    nodes = list(zip(ids, (int(v) for v in c["ndof"])))
    ne = len(c["parent"])
    elems, children = [], []
    for e in range(ne):
        nl = [ids[k] for k in c["enode"][c["eptr"][e]:c["eptr"][e + 1]]]
        if nl:
            elems.append(nl)
        children.append([int(v) for v in c["cidx"][c["cptr"][e]:c["cptr"][e + 1]]])
    return MeshSpec(nodes, elems, [int(v) for v in c["parent"]], children)


def lca_lock(spec):
    """Lock results by the lowest-common-ancestor formulation: a node is a
    factor node of the LCA of the leaves that contain it and a schur node of
    every element on the paths from those leaves to (not including) the LCA;
    offsets are numbered generation 0 first, element id order, node creation
    order.  Equivalent to the reference's owner-count process for
    default-kind nodes; written independently of it."""
    ne = spec.n_elem
    gen = [0] * ne

    def g(e):
        if not spec.children[e]:
            gen[e] = 0
            return 0
        gen[e] = min(g(c) for c in spec.children[e]) - 1
        return gen[e]

    for e in range(ne):
        if spec.parent[e] < 0:
            g(e)
    depth = [0] * ne
    for e in range(ne):
        k, x = 0, e
        while spec.parent[x] >= 0:
            x = spec.parent[x]
            k += 1
        depth[e] = k
    owners = {}
    for e, nl in enumerate(spec.elems):
        for n in nl:
            owners.setdefault(n, []).append(e)
    factor = [set() for _ in range(ne)]
    schur = [set() for _ in range(ne)]
    for n, ls in owners.items():
        paths = []
        for leaf in ls:
            p, x = [leaf], leaf
            while spec.parent[x] >= 0:
                x = spec.parent[x]
                p.append(x)
            paths.append(p)
        common = set(paths[0])
        for p in paths[1:]:
            common &= set(p)
        lca = max(common, key=lambda e: depth[e])
        factor[lca].add(n)
        for p in paths:
            for x in p:
                if x == lca:
                    break
                schur[x].add(n)
    order = sorted(range(ne), key=lambda e: (-gen[e], e))
    off, go = {}, 0
    for e in order:
        for n in sorted(factor[e], key=lambda n: spec.seq[n]):
            off[n] = go
            go += spec.ndof[n]
    res = []
    for e in range(ne):
        f = sorted((n for n in factor[e] if spec.ndof[n]), key=lambda n: off[n])
        s = sorted((n for n in schur[e] if spec.ndof[n]), key=lambda n: off[n])
        fo, so, acc = [], [], 0
        for n in f:
            fo.append(acc)
            acc += spec.ndof[n]
        fs, acc = acc, 0
        for n in s:
            so.append(acc)
            acc += spec.ndof[n]
        res.append(dict(generation=gen[e], fs=fs, ss=acc, factor=f, factor_off=fo, schur=s, schur_off=so))
    return res, off, go


def front_index_map(el_parent, el_child):
    """Unified parent index of every child schur dof, by direct lookup of
    each node in the parent's factor / schur lists."""
    pos = {}
    for n, o in zip(el_parent["factor"], el_parent["factor_off"]):
        pos[n] = o
    for n, o in zip(el_parent["schur"], el_parent["schur_off"]):
        pos[n] = el_parent["fs"] + o
    return pos


def load_q2():
    return np.load(os.path.join(GOLDEN, "q2_mesh.npz"))
