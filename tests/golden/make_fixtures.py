"""Converts the reference's own test data (sloam/src/tests/aux/) into compact
fixtures under tests/golden/.  Run in a container that has the reference
mounted at /root/reference; the outputs are committed, the reference is not
needed afterwards.

  {still,moving}_tree_t{0,1}.npz
                          organised 2048 x 64 tree clouds (ASCII PCD): the
                          flat indices and xyz of their finite points
  {still,moving}_landmarks_t{0,1}.npz
                          boost text_iarchive (v17) of
                          std::vector<std::vector<TreeVertex>>: per vertex
                          treeId, beam, prevVertexSize, radius, isValid,
                          coords xyz and its points xyz (ragged, flattened)
  still_ground_t{0,1}.npz ASCII PCD ground clouds: xyz float32
"""
import os
import sys

import numpy as np

AUX = "/root/reference/sloam/src/tests/aux"
OUT = os.path.dirname(os.path.abspath(__file__))


def read_pcd(path):
    with open(path) as f:
        header = {}
        while True:
            line = f.readline()
            key = line.split()[0]
            header[key] = line.split()[1:]
            if key == "DATA":
                break
        data = np.loadtxt(f, dtype=np.float64, ndmin=2)
    w, h = int(header["WIDTH"][0]), int(header["HEIGHT"][0])
    assert header["DATA"][0] == "ascii"
    assert data.shape[0] == w * h, (data.shape, w, h)
    return data, w, h


class Tokens:
    def __init__(self, text):
        self.t = text.split()
        self.i = 0

    def next(self):
        v = self.t[self.i]
        self.i += 1
        return v

    def int(self):
        return int(self.next())

    def float(self):
        return float(self.next())


def read_landmarks(path):
    """boost::archive::text_iarchive of vector<vector<TreeVertex>>; class
    headers (tracking level, version) appear at the first occurrence of each
    serialised type only."""
    tk = Tokens(open(path).read())
    assert tk.int() == 22 and tk.next() == "serialization::archive"
    tk.int()  # archive version (17)
    tk.int(); tk.int()  # outer vector class header
    n_trees = tk.int()
    tk.int()  # item version
    first_inner = first_vtx = first_pt = first_pts = True
    trees = []
    for _ in range(n_trees):
        if first_inner:
            tk.int(); tk.int()
            first_inner = False
        n_v = tk.int()
        tk.int()  # item version
        verts = []
        for _ in range(n_v):
            if first_vtx:
                tk.int(); tk.int()
                first_vtx = False
            tree_id = tk.int()
            beam = tk.int()
            prev = tk.int()
            radius = tk.float()
            valid = tk.int()
            if first_pt:
                tk.int(); tk.int()
                first_pt = False
            coords = [tk.float(), tk.float(), tk.float()]
            if first_pts:
                tk.int(); tk.int()
                first_pts = False
            n_p = tk.int()
            tk.int()  # item version
            pts = [[tk.float(), tk.float(), tk.float()] for _ in range(n_p)]
            verts.append((tree_id, beam, prev, radius, valid, coords, pts))
        trees.append(verts)
    assert tk.i == len(tk.t), (tk.i, len(tk.t))
    return trees


def save_landmarks(trees, out):
    tree_off, vert_off = [0], [0]
    vi, pts = [], []
    for verts in trees:
        for (tid, beam, prev, rad, valid, coords, p) in verts:
            vi.append([tid, beam, prev, valid])
            pts.extend(p)
            vert_off.append(len(pts))
        tree_off.append(len(vert_off) - 1)
    vf = [[v[3]] + v[5] for t in trees for v in t]
    np.savez_compressed(out, tree_offsets=np.array(tree_off, np.int32),
                        vertex_offsets=np.array(vert_off, np.int32),
                        vertex_int=np.array(vi, np.int32),
                        vertex_radius_coords=np.array(vf, np.float64),
                        points=np.array(pts, np.float32))


def main():
    if not os.path.isdir(AUX):
        print("reference test data not found", file=sys.stderr)
        return 1
    for name in ("still_tree_t0", "still_tree_t1", "moving_tree_t0", "moving_tree_t1"):
        data, w, h = read_pcd(os.path.join(AUX, name + ".pcd"))
        fin = np.isfinite(data[:, :3]).all(1)
        np.savez_compressed(os.path.join(OUT, name + ".npz"), width=w, height=h,
                            index=np.nonzero(fin)[0].astype(np.int32),
                            xyz=data[fin, :3].astype(np.float32))
    for pre in ("still", "moving"):
        for t in ("t0", "t1"):
            trees = read_landmarks(os.path.join(AUX, f"{pre}_landmarks_{t}"))
            save_landmarks(trees, os.path.join(OUT, f"{pre}_landmarks_{t}.npz"))
    for t in ("t0", "t1"):
        data, w, h = read_pcd(os.path.join(AUX, f"still_ground_{t}.pcd"))
        np.savez_compressed(os.path.join(OUT, f"still_ground_{t}.npz"),
                            xyz=data[:, :3].astype(np.float32))
    return 0


if __name__ == "__main__":
    sys.exit(main())
