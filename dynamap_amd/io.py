"""Host-side I/O around the mapping path: TUM RGB-D sequences, trajectory
files and their ATE / RPE evaluation, and mesh files (OBJ read/write, PLY
write).

Reference (paths relative to the reference root):
  include/utils/utils_Dataset.cpp:58-247   dataSet: association file, image
                                           loading, ground-truth lines
  include/builder/builder.cu:222-262       v2tRef (TUM line -> pose),
                                           writePoseToTxt (pose -> TUM line)
  include/geometry/mesh/mesh_IO.h:32-90    Mesh_IO readOBJ / saveOBJ /
                                           updateVerticesNormals
Differences: depth PNGs are read as 16-bit (the reference opens them with
IMREAD_COLOR, utils_Dataset.cpp:233-235, which truncates 16-bit depth to
8 bits); without an association file, frames are associated by nearest
timestamp (max 20 ms), as TUM's associate.py does; ground truth is taken at
the nearest timestamp instead of a fixed line offset.
"""
import os

import numpy as np


# ---------------------------------------------------------------- poses
def quat_to_rot(qx, qy, qz, qw):
    """Rotation matrix of a (normalised) quaternion."""
    q = np.array([qw, qx, qy, qz], np.float64)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def rot_to_quat(R):
    """(qx, qy, qz, qw) of a rotation matrix, qw >= 0."""
    R = np.asarray(R, np.float64)
    tr = R[0, 0] + R[1, 1] + R[2, 2]
    if tr > 0:
        s = np.sqrt(tr + 1.0) * 2
        w, x, y, z = 0.25 * s, (R[2, 1] - R[1, 2]) / s, (R[0, 2] - R[2, 0]) / s, (R[1, 0] - R[0, 1]) / s
    elif R[0, 0] > R[1, 1] and R[0, 0] > R[2, 2]:
        s = np.sqrt(1.0 + R[0, 0] - R[1, 1] - R[2, 2]) * 2
        w, x, y, z = (R[2, 1] - R[1, 2]) / s, 0.25 * s, (R[0, 1] + R[1, 0]) / s, (R[0, 2] + R[2, 0]) / s
    elif R[1, 1] > R[2, 2]:
        s = np.sqrt(1.0 + R[1, 1] - R[0, 0] - R[2, 2]) * 2
        w, x, y, z = (R[0, 2] - R[2, 0]) / s, (R[0, 1] + R[1, 0]) / s, 0.25 * s, (R[1, 2] + R[2, 1]) / s
    else:
        s = np.sqrt(1.0 + R[2, 2] - R[0, 0] - R[1, 1]) * 2
        w, x, y, z = (R[1, 0] - R[0, 1]) / s, (R[0, 2] + R[2, 0]) / s, (R[1, 2] + R[2, 1]) / s, 0.25 * s
    q = np.array([x, y, z, w])
    return q if w >= 0 else -q


def tum_line_to_pose(vals):
    """builder::v2tRef (builder.cu:222-242): [t, tx, ty, tz, qx, qy, qz, qw] -> 4x4."""
    T = np.eye(4)
    T[:3, :3] = quat_to_rot(*vals[4:8])
    T[:3, 3] = vals[1:4]
    return T


def read_trajectory(path):
    """TUM trajectory file -> (timestamps (n,), poses (n,4,4)); '#' lines skipped."""
    stamps, poses = [], []
    with open(path) as f:
        for line in f:
            line = line.strip()
            if not line or line[0] == "#":
                continue
            v = [float(x) for x in line.replace(",", " ").split()]
            stamps.append(v[0])
            poses.append(tum_line_to_pose(v))
    return np.array(stamps), np.array(poses).reshape(-1, 4, 4)


def write_trajectory(path, stamps, poses, append=False):
    """builder::writePoseToTxt (builder.cu:243-262): "t tx ty tz qx qy qz qw",
    fixed with 4 decimals, one line per pose."""
    with open(path, "a" if append else "w") as f:
        for t, T in zip(stamps, poses):
            q = rot_to_quat(T[:3, :3])
            f.write(f"{t:.4f} {T[0, 3]:.4f} {T[1, 3]:.4f} {T[2, 3]:.4f} "
                    f"{q[0]:.4f} {q[1]:.4f} {q[2]:.4f} {q[3]:.4f}\n")


# ---------------------------------------------------------------- evaluation
def align_rigid(A, B):
    """R, t minimising sum |R a_i + t - b_i|^2 (Horn / Umeyama without scale)."""
    A = np.asarray(A, np.float64)
    B = np.asarray(B, np.float64)
    ma, mb = A.mean(0), B.mean(0)
    H = (A - ma).T @ (B - mb)
    U, _, Vt = np.linalg.svd(H)
    D = np.eye(3)
    D[2, 2] = np.sign(np.linalg.det(Vt.T @ U.T))
    R = Vt.T @ D @ U.T
    return R, mb - R @ ma


def ate(est, gt, align=True):
    """Absolute trajectory error: RMSE of the translations after rigid
    alignment of est to gt (est, gt: (n,4,4) or (n,3))."""
    pe = est[:, :3, 3] if est.ndim == 3 else est
    pg = gt[:, :3, 3] if gt.ndim == 3 else gt
    if align:
        R, t = align_rigid(pe, pg)
        pe = pe @ R.T + t
    return float(np.sqrt(np.mean(np.sum((pe - pg) ** 2, 1))))


def rpe(est, gt, delta=1):
    """Relative pose error over `delta` frames: (translational RMSE, mean
    rotation angle in radians) of inv(gt_i^-1 gt_j) est_i^-1 est_j."""
    et, er = [], []
    for i in range(len(est) - delta):
        j = i + delta
        dg = np.linalg.inv(gt[i]) @ gt[j]
        de = np.linalg.inv(est[i]) @ est[j]
        E = np.linalg.inv(dg) @ de
        et.append(np.linalg.norm(E[:3, 3]))
        er.append(np.arccos(np.clip((np.trace(E[:3, :3]) - 1) / 2, -1, 1)))
    et = np.array(et)
    return float(np.sqrt(np.mean(et ** 2))) if len(et) else 0.0, float(np.mean(er)) if len(er) else 0.0


# ---------------------------------------------------------------- TUM RGB-D
def _read_list(path):
    out = []
    with open(path) as f:
        for line in f:
            line = line.strip()
            if not line or line[0] == "#":
                continue
            w = line.split()
            out.append((float(w[0]), w[1:]))
    return out


def _nearest(stamps, t):
    k = int(np.searchsorted(stamps, t))
    cand = [c for c in (k - 1, k) if 0 <= c < len(stamps)]
    return min(cand, key=lambda c: abs(stamps[c] - t))


class TumDataset:
    """A TUM RGB-D sequence (rgb.txt, depth.txt, groundtruth.txt, optional
    associations.txt) as the reference's dataSet (utils_Dataset.cpp).

    frame(i) -> (depth float32 metres (rows, cols), rgb uint8 (rows, cols, 3),
    ground-truth pose (4,4) or None, timestamp)."""

    def __init__(self, root, depth_scale=5000.0, association="associations.txt", max_difference=0.02):
        self.root = root
        self.depth_scale = float(depth_scale)
        assoc = os.path.join(root, association)
        pairs = []
        if os.path.exists(assoc):
            for t, rest in _read_list(assoc):
                # "t_a file_a t_b file_b": the depth entry is the one under depth/
                fa, tb, fb = rest[0], float(rest[1]), rest[2]
                if "depth" in fa:
                    pairs.append((t, fa, tb, fb))
                else:
                    pairs.append((tb, fb, t, fa))
        else:
            depth = _read_list(os.path.join(root, "depth.txt"))
            rgb = _read_list(os.path.join(root, "rgb.txt"))
            rs = np.array([t for t, _ in rgb])
            for t, rest in depth:
                if not len(rs):
                    break
                k = _nearest(rs, t)
                if abs(rs[k] - t) <= max_difference:
                    pairs.append((t, rest[0], rgb[k][0], rgb[k][1][0]))
        self.pairs = pairs
        gt_path = os.path.join(root, "groundtruth.txt")
        if os.path.exists(gt_path):
            self.gt_stamps, self.gt_poses = read_trajectory(gt_path)
        else:
            self.gt_stamps, self.gt_poses = None, None

    def __len__(self):
        return len(self.pairs)

    def timestamps(self):
        return np.array([p[0] for p in self.pairs])

    def groundtruth(self, t):
        if self.gt_stamps is None or not len(self.gt_stamps):
            return None
        return self.gt_poses[_nearest(self.gt_stamps, t)].copy()

    def frame(self, i):
        from PIL import Image
        td, fd, tr, fr = self.pairs[i]
        d = np.asarray(Image.open(os.path.join(self.root, fd)), dtype=np.float32) / self.depth_scale
        rgb = np.asarray(Image.open(os.path.join(self.root, fr)).convert("RGB"), dtype=np.uint8)
        return np.ascontiguousarray(d), np.ascontiguousarray(rgb), self.groundtruth(td), td


def write_tum_sequence(root, depth, rgb, poses, stamps=None, depth_scale=5000.0):
    """Write frames in the TUM layout (rgb/, depth/ 16-bit PNG, rgb.txt,
    depth.txt, groundtruth.txt, associations.txt)."""
    from PIL import Image
    os.makedirs(os.path.join(root, "rgb"), exist_ok=True)
    os.makedirs(os.path.join(root, "depth"), exist_ok=True)
    n = len(depth)
    stamps = np.arange(n) / 30.0 + 1.0 if stamps is None else stamps
    with open(os.path.join(root, "rgb.txt"), "w") as fr, open(os.path.join(root, "depth.txt"), "w") as fd, \
            open(os.path.join(root, "associations.txt"), "w") as fa:
        fr.write("# timestamp filename\n")
        fd.write("# timestamp filename\n")
        for k in range(n):
            t = stamps[k]
            name = f"{t:.6f}.png"
            dq = np.clip(np.round(np.asarray(depth[k], np.float64) * depth_scale), 0, 65535).astype(np.uint16)
            Image.fromarray(dq).save(os.path.join(root, "depth", name))
            Image.fromarray(np.asarray(rgb[k], np.uint8)).save(os.path.join(root, "rgb", name))
            fr.write(f"{t:.6f} rgb/{name}\n")
            fd.write(f"{t:.6f} depth/{name}\n")
            fa.write(f"{t:.6f} rgb/{name} {t:.6f} depth/{name}\n")
    with open(os.path.join(root, "groundtruth.txt"), "w") as f:
        f.write("# timestamp tx ty tz qx qy qz qw\n")
        for t, T in zip(stamps, poses):
            q = rot_to_quat(T[:3, :3])
            f.write(f"{t:.6f} " + " ".join(f"{x:.9f}" for x in (*T[:3, 3], *q)) + "\n")
    return stamps


# ---------------------------------------------------------------- meshes
def vertex_normals(V, F):
    """Area-weighted vertex normals (Mesh_IO::updateVerticesNormals)."""
    V = np.asarray(V, np.float64)
    n = np.cross(V[F[:, 1]] - V[F[:, 0]], V[F[:, 2]] - V[F[:, 0]])
    out = np.zeros_like(V)
    for k in range(3):
        np.add.at(out, F[:, k], n)
    ln = np.linalg.norm(out, axis=1, keepdims=True)
    return out / np.where(ln > 0, ln, 1)


def save_obj(path, V, F, colors=None, normals=None):
    """Indexed OBJ (Mesh_IO::saveOBJ): "v x y z [r g b]", optional "vn",
    faces 1-based ("f a//a b//b c//c" with normals)."""
    with open(path, "w") as f:
        for k, p in enumerate(V):
            if colors is not None:
                c = colors[k]
                f.write(f"v {p[0]:g} {p[1]:g} {p[2]:g} {c[0]:g} {c[1]:g} {c[2]:g}\n")
            else:
                f.write(f"v {p[0]:g} {p[1]:g} {p[2]:g}\n")
        if normals is not None:
            for n in normals:
                f.write(f"vn {n[0]:g} {n[1]:g} {n[2]:g}\n")
            for a, b, c in F + 1:
                f.write(f"f {a}//{a} {b}//{b} {c}//{c}\n")
        else:
            for a, b, c in F + 1:
                f.write(f"f {a} {b} {c}\n")


def read_obj(path):
    """Mesh_IO::readOBJ: vertices (and per-vertex colours when present),
    triangles (polygons fanned), 0-based."""
    V, C, F = [], [], []
    with open(path) as f:
        for line in f:
            w = line.split()
            if not w:
                continue
            if w[0] == "v":
                V.append([float(x) for x in w[1:4]])
                if len(w) >= 7:
                    C.append([float(x) for x in w[4:7]])
            elif w[0] == "f":
                idx = [int(t.split("/")[0]) - 1 for t in w[1:]]
                for k in range(1, len(idx) - 1):
                    F.append([idx[0], idx[k], idx[k + 1]])
    V = np.array(V, np.float32).reshape(-1, 3)
    C = np.array(C, np.float32).reshape(-1, 3) if len(C) == len(V) and len(C) else None
    return V, np.array(F, np.int64).reshape(-1, 3), C


def save_ply(path, V, F, colors=None):
    """Binary little-endian PLY: float xyz, optional uchar rgb (colours in
    [0, 1]), int32 triangle lists."""
    V = np.asarray(V, np.float32)
    F = np.asarray(F, np.int32)
    with open(path, "wb") as f:
        hdr = ["ply", "format binary_little_endian 1.0", f"element vertex {len(V)}",
               "property float x", "property float y", "property float z"]
        if colors is not None:
            hdr += ["property uchar red", "property uchar green", "property uchar blue"]
        hdr += [f"element face {len(F)}", "property list uchar int vertex_indices", "end_header"]
        f.write(("\n".join(hdr) + "\n").encode())
        if colors is not None:
            c8 = np.clip(np.round(np.asarray(colors, np.float64) * 255.0), 0, 255).astype(np.uint8)
            rec = np.zeros(len(V), dtype=[("p", "<f4", 3), ("c", "u1", 3)])
            rec["p"], rec["c"] = V, c8
        else:
            rec = np.zeros(len(V), dtype=[("p", "<f4", 3)])
            rec["p"] = V
        f.write(rec.tobytes())
        frec = np.zeros(len(F), dtype=[("n", "u1"), ("i", "<i4", 3)])
        frec["n"], frec["i"] = 3, F
        f.write(frec.tobytes())


def read_ply(path):
    """Reader for the files save_ply writes (binary little-endian)."""
    with open(path, "rb") as f:
        nv = nf = 0
        has_c = False
        while True:
            line = f.readline().decode().strip()
            if line.startswith("element vertex"):
                nv = int(line.split()[-1])
            elif line.startswith("element face"):
                nf = int(line.split()[-1])
            elif line.startswith("property uchar red"):
                has_c = True
            elif line == "end_header":
                break
        dt = [("p", "<f4", 3)] + ([("c", "u1", 3)] if has_c else [])
        rec = np.frombuffer(f.read(np.dtype(dt).itemsize * nv), dtype=dt)
        frec = np.frombuffer(f.read(np.dtype([("n", "u1"), ("i", "<i4", 3)]).itemsize * nf),
                             dtype=[("n", "u1"), ("i", "<i4", 3)])
    return rec["p"].copy(), frec["i"].astype(np.int64), (rec["c"].copy() if has_c else None)
