"""Pure-Python restatement of clean_hex_mesh's flag passes, for small meshes.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the library's hexclean_*
calls. The product path never imports this module.

It keeps the reference's Mesh as plain lists (faces, edges and their neighbour lists, built
the way build_connectivity builds them for hexes, global_functions.cpp:121-183 and 225-262)
and walks them in the reference's order:
  reorder_hex_mesh          global_functions.cpp:2199-2229
  tagging_uneven_element    grid_meshing/grid_hex_meshing.cpp:1983-2004
  re_indexing_connectivity  global_functions.cpp:664-698
  clean_non_manifold_ve     grid_meshing/grid_hex_meshing.cpp:2006-2080
  drop_small_pieces         grid_meshing/grid_hex_meshing.cpp:2081-2124
  clean_hex_mesh's order    grid_meshing/grid_hex_meshing.cpp:1954-1964
Pinned by hand-made meshes whose cleaned flags follow from the definitions (a spike, a pit,
pieces joined along an edge or at a vertex, separate pieces) in tests/test_hexclean.py.
"""
HEX_FACE = ((0, 1, 2, 3), (4, 7, 6, 5), (0, 4, 5, 1), (0, 3, 7, 4), (3, 2, 6, 7), (1, 5, 6, 2))
HEX_TET = ((0, 3, 4, 1), (1, 0, 5, 2), (2, 1, 6, 3), (3, 2, 7, 0), (4, 7, 5, 0), (5, 4, 6, 1), (6, 5, 7, 2),
           (7, 6, 4, 3))


class Mesh:
    """Vs/Es/Fs/Hs of a hex mesh with local vertex ids 0..nv-1."""

    def __init__(self, nv, hexes):
        self.nv = nv
        self.Hvs = [list(h) for h in hexes]
        temp = []
        total = []
        for i, h in enumerate(self.Hvs):
            for j in range(6):
                vs = [h[k] for k in HEX_FACE[j]]
                total.append(vs)
                temp.append((tuple(sorted(vs)), 6 * i + j, i, j))
        temp.sort()
        self.Hfs = [[None] * 6 for _ in self.Hvs]
        self.Fvs, self.Fbnd = [], []
        for n, (key, fid6, i, j) in enumerate(temp):
            if n == 0 or key != temp[n - 1][0]:
                self.Fvs.append(total[fid6])
                self.Fbnd.append(True)
            else:
                self.Fbnd[-1] = False
            self.Hfs[i][j] = len(self.Fvs) - 1
        etemp = []
        for f, vs in enumerate(self.Fvs):
            for j in range(4):
                a, b = vs[j], vs[(j + 1) % 4]
                etemp.append((min(a, b), max(a, b), f, j))
        etemp.sort()
        self.Fes = [[None] * 4 for _ in self.Fvs]
        ne = 0
        for n, (a, b, f, j) in enumerate(etemp):
            if n == 0 or (a, b) != etemp[n - 1][:2]:
                ne += 1
            self.Fes[f][j] = ne - 1
        self.Ebnd = [False] * ne
        self.Vbnd = [False] * nv
        for f, vs in enumerate(self.Fvs):
            if self.Fbnd[f]:
                for j in range(4):
                    self.Ebnd[self.Fes[f][j]] = True
                    self.Vbnd[vs[j]] = True
        self.Fhs = [[] for _ in self.Fvs]
        for i in range(len(self.Hvs)):
            for j in range(6):
                self.Fhs[self.Hfs[i][j]].append(i)
        self.Efs = [[] for _ in range(ne)]
        self.Vfs = [[] for _ in range(nv)]
        for f in range(len(self.Fvs)):
            for j in range(4):
                self.Efs[self.Fes[f][j]].append(f)
                self.Vfs[self.Fvs[f][j]].append(f)


def reindex(H, flag):
    """re_indexing_connectivity: the flagged hexes with their vertices renumbered in order."""
    used = sorted({v for h, f in zip(H, flag) if f for v in h})
    vmap = {v: i for i, v in enumerate(used)}
    hmap = [i for i, f in enumerate(flag) if f]
    return Mesh(len(used), [[vmap[v] for v in H[i]] for i in hmap]), hmap


def _det(a, b, c, d):
    m = [[b[r] - a[r], c[r] - a[r], d[r] - a[r]] for r in range(3)]
    h = lambda x, y, z: m[0][x] * (m[1][y] * m[2][z] - m[1][z] * m[2][y])  # noqa: E731
    return h(0, 1, 2) - h(1, 0, 2) + h(2, 0, 1)


def reorder_hex_mesh(V, H):
    out, n = [], 0
    for h in H:
        vol = 0.0
        for t in HEX_TET:
            vol += _det(*(V[h[k]] for k in t))
        if vol < 0:
            h = [h[3], h[2], h[1], h[0], h[7], h[6], h[5], h[4]]
            n += 1
        out.append(list(h))
    return out, n


def tagging_uneven_element(nv, H, flag):
    flag = list(flag)
    m = Mesh(nv, H)
    for i in range(len(H)):
        bn = inn = 0
        for fid in m.Hfs[i]:
            nh = m.Fhs[fid]
            if len(nh) == 1:
                bn += 1
            elif flag[i]:
                if not flag[nh[0]] or not flag[nh[1]]:
                    bn += 1
                else:
                    inn += 1
            else:
                if flag[nh[0]] or flag[nh[1]]:
                    bn += 1
                else:
                    inn += 1
        if bn == 5 and inn == 1:
            flag[i] = not flag[i]
    return flag


def clean_non_manifold_ve(nv, H, flag):
    flag = list(flag)
    m, hmap = reindex(H, flag)
    while True:
        changed = False
        v_tag = [False] * m.nv
        for v in range(m.nv):
            if not m.Vbnd[v] or v_tag[v]:
                continue
            v_tag[v] = True
            bfs = [f for f in m.Vfs[v] if m.Fbnd[f]]
            tag = [False] * len(bfs)
            pool = [bfs[0]]
            bad = False
            while True:
                nxt = []
                for fid in pool:
                    k = bfs.index(fid)
                    if tag[k]:
                        continue
                    tag[k] = True
                    for eid in m.Fes[fid]:
                        for nf in m.Efs[eid]:
                            if nf in bfs and not tag[bfs.index(nf)]:
                                nxt.append(nf)
                if nxt:
                    pool = nxt
                elif not all(tag):
                    bad = True
                    break
                else:
                    break
            if bad:
                for k, f in enumerate(bfs):
                    if not tag[k]:
                        nh = m.Fhs[f][0]
                        flag[hmap[nh]] = False
                        for u in m.Hvs[nh]:
                            v_tag[u] = True
                changed = True
        if not changed:
            e_tag = [False] * len(m.Ebnd)
            for e in range(len(m.Ebnd)):
                if not m.Ebnd[e] or e_tag[e]:
                    continue
                e_tag[e] = True
                bfs = [f for f in m.Efs[e] if m.Fbnd[f]]
                if len(bfs) != 2:
                    changed = True
                    nh = m.Fhs[bfs[0]][0]
                    flag[hmap[nh]] = False
                    for fid in m.Hfs[nh]:
                        for eid in m.Fes[fid]:
                            e_tag[eid] = True
            if not changed:
                break
        m, hmap = reindex(H, flag)
    return flag


def drop_small_pieces(nv, H, flag):
    flag = list(flag)
    m, hmap = reindex(H, flag)
    n = len(hmap)
    if n == 0:
        return flag
    seen = [False] * n
    groups = []
    group = [0]
    while True:
        i = 0
        while i < len(group):
            hid = group[i]
            i += 1
            if seen[hid]:
                continue
            seen[hid] = True
            for fid in m.Hfs[hid]:
                for nh in m.Fhs[fid]:
                    if nh != hid and not seen[nh]:
                        group.append(nh)
        groups.append(group)
        if not all(seen):
            group = [seen.index(False)]
        else:
            break
    if len(groups) > 1:
        which = 0
        for g in range(1, len(groups)):
            if len(groups[g]) > len(groups[which]):
                which = g
        flag = [False] * len(flag)
        for hid in groups[which]:
            flag[hmap[hid]] = True
    return flag


def clean_flags(nv, H, flag):
    flag = tagging_uneven_element(nv, H, flag)
    if not any(flag):
        return flag
    flag = clean_non_manifold_ve(nv, H, flag)
    return drop_small_pieces(nv, H, flag)
