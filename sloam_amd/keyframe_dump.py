"""Reader of sloam_amd/keyframe_bench --dump files: per RunSloam call of the first pass,
the estimator's input as the keyframe runner hands it over, its state before
the call and its output (record layout: sloam_amd/bench/keyframe_bench.cpp,
DumpWriter)."""
import numpy as np


class _Cursor:
    def __init__(self, buf):
        self.b = buf
        self.o = 0

    def take(self, dtype, n=1):
        dt = np.dtype(dtype)
        a = np.frombuffer(self.b, dtype=dt, count=n, offset=self.o)
        self.o += dt.itemsize * n
        return a

    def i32(self):
        return int(self.take("<i4")[0])

    def done(self):
        return self.o >= len(self.b)


def read(path):
    c = _Cursor(open(path, "rb").read())
    recs = []
    while not c.done():
        r = {"keyframe": c.i32(), "first": bool(c.i32())}
        r["pose"] = c.take("<f8", 7).copy()
        nc, nf = c.i32(), c.i32()
        r["nf"] = nf
        cm, cv, cf = [], [], []
        for _ in range(nc):
            cm.append(c.take("<f8", 7))
            cv.append(c.i32())
            cf.append(c.take("<f4", 4 * nf))
        r["cell_model"] = np.array(cm).reshape(nc, 7)
        r["cell_valid"] = np.array(cv, np.int32)
        r["cell_feat"] = np.array(cf, np.float32).reshape(nc, nf, 4)
        nt = c.i32()
        tree_off, vtx_int, vtx_f64, vtx_pt_off, pts = [0], [], [], [0], []
        for _ in range(nt):
            nv = c.i32()
            for _ in range(nv):
                vid, valid = c.i32(), c.i32()
                vtx_int.append((vid, valid))
                vtx_f64.append(c.take("<f8", 4))
                npts = c.i32()
                pts.append(c.take("<f4", 4 * npts).reshape(npts, 4))
                vtx_pt_off.append(vtx_pt_off[-1] + npts)
            tree_off.append(tree_off[-1] + nv)
        r["tree_off"] = np.array(tree_off, np.int32)
        r["vtx_int"] = np.array(vtx_int, np.int32).reshape(-1, 2)
        r["vtx_f64"] = np.array(vtx_f64).reshape(-1, 4)
        r["vtx_pt_off"] = np.array(vtx_pt_off, np.int32)
        r["pts"] = np.concatenate(pts) if pts else np.zeros((0, 4), np.float32)
        nl = c.i32()
        li, lc = [], []
        for _ in range(nl):
            li.append((c.i32(), c.i32()))
            lc.append(c.take("<f4", 6))
        r["line_int"] = np.array(li, np.int32).reshape(-1, 2)
        r["line_coeff"] = np.array(lc, np.float32).reshape(-1, 6)
        ncf = c.i32()
        ok, fit = [], []
        for _ in range(ncf):
            ok.append(c.i32())
            fit.append(c.take("<f8", 7))
        r["cyl_ok"] = np.array(ok, np.int32)
        r["cyl_fit"] = np.array(fit).reshape(-1, 7)
        nm = c.i32()
        r["map"] = c.take("<f8", 7 * nm).reshape(nm, 7).copy()
        npv = c.i32()
        r["prev_planes"] = c.take("<f8", 7 * npv).reshape(npv, 7).copy()
        # output of the call
        o = {"ok": bool(c.i32()), "tree_step": bool(c.i32()), "ground_step": bool(c.i32()),
             "tree_matches": c.i32(), "plane_matches": c.i32(), "tree_iters": c.i32(),
             "ground_iters": c.i32()}
        o["pose"] = c.take("<f8", 7).copy()
        o["delta"] = c.take("<f8", 7).copy()
        ntm = c.i32()
        ids, tm, matches = [], [], []
        for _ in range(ntm):
            ids.append(c.i32())
            tm.append(c.take("<f8", 7))
            matches.append(c.i32())
        o["tm_id"] = np.array(ids, np.int32)
        o["tm"] = np.array(tm).reshape(ntm, 7)
        o["matches"] = np.array(matches, np.int32)
        npa = c.i32()
        o["planes"] = c.take("<f8", 7 * npa).reshape(npa, 7).copy()
        r["out"] = o
        recs.append(r)
    return recs


def quat_to_R(q):
    """q = x, y, z, w."""
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def unproject_trees(tm, pose):
    """Tree models [k, 7] (root, ray, radius) in the map frame of `pose` back
    in the sensor frame."""
    R = quat_to_R(pose[3:])
    t = pose[:3]
    out = tm.copy()
    out[:, :3] = (tm[:, :3] - t) @ R
    out[:, 3:6] = tm[:, 3:6] @ R
    return out


def unproject_planes(pl, pose):
    """Plane models [k, 7] (plane[4], centroid[3]) in the map frame back in
    the sensor frame."""
    R = quat_to_R(pose[3:])
    t = pose[:3]
    out = pl.copy()
    n = pl[:, :3]
    out[:, :3] = n @ R
    out[:, 3] = pl[:, 3] + n @ t
    out[:, 4:7] = (pl[:, 4:7] - t) @ R
    return out
