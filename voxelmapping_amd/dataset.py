"""TUM-format RGB-D sequence loader (DatasetLoader, example/dataset_loader.h:12-27,
example/dataset_loader.cpp:6-159).

A sequence directory holds:
  * calibration.txt - first line ``fx fy cx cy`` (dataset_loader.cpp:143-160);
  * associated.txt  - lines ``t_rgb rgb_path t_depth depth_path``; lines that
    are empty or start with '#' are skipped; the frame time is t_rgb
    (dataset_loader.cpp:42-70);
  * groundtruth.txt - lines ``t tx ty tz qx qy qz qw`` (camera-to-world;
    the quaternion is normalised), matched to every frame by the smallest
    |t_gt - t_frame| (first minimum wins, differences of 1000 s or more never
    match) (dataset_loader.cpp:72-122).
Depth PNGs are 16-bit with 5000 units per metre: the example converts them
with ``convertTo(CV_32FC1, 1 / 5000.0)`` (pangolin_display_3d.cpp:32-33), so
``depth_scale`` is 1 / 5000 and the raw uint16 image can be fused directly
with ``DeviceMap.fuse(depth_u16, ..., depth_scale=1/5000)``.
"""
from __future__ import annotations

import os

import numpy as np

TUM_DEPTH_SCALE = 1.0 / 5000.0


def _read_png(path: str, bgr: bool = False) -> np.ndarray:
    from PIL import Image  # decoding only; PIL is the image library present here
    with Image.open(path) as im:
        if im.mode in ("I;16", "I;16B", "I;16L"):
            return np.array(im, dtype=np.uint16)
        if im.mode == "I":
            return np.array(im).astype(np.uint16)
        if im.mode in ("RGB", "RGBA", "L"):
            c = np.array(im.convert("RGB"), dtype=np.uint8)
            return np.ascontiguousarray(c[..., ::-1]) if bgr else c
        raise ValueError(f"{path}: unsupported PNG mode {im.mode}")


def quaternion_pose(tx, ty, tz, qx, qy, qz, qw) -> np.ndarray:
    """4x4 camera-to-world from a translation and an (x, y, z, w) quaternion,
    normalised first (Eigen Quaterniond(qw, qx, qy, qz).normalize(),
    toRotationMatrix; dataset_loader.cpp:88-96)."""
    q = np.array([qw, qx, qy, qz], np.float64)
    w, x, y, z = q / np.linalg.norm(q)
    T = np.eye(4)
    T[:3, :3] = [[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                 [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                 [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]]
    T[:3, 3] = [tx, ty, tz]
    return T


def _records(path: str):
    with open(path) as f:
        for line in f:
            line = line.rstrip("\n")
            if line and not line.startswith("#"):
                yield line.split()


class DatasetLoader:
    """Mirror of vmap::DatasetLoader. ``GetNext()`` returns
    ``(depth_u16 HxW, color HxWx3 uint8, time, camToWorld 4x4 float64)`` or
    None at the end; the pose is None when no ground truth was loaded.

    The colour comes in the reference's channel order by default: it reads
    images with ``cv::imread(path, -1)`` (dataset_loader.cpp:131-132), which
    gives blue, green, red; ``color_order="rgb"`` gives red, green, blue. An
    alpha channel is dropped and a grey image is repeated into three channels
    (the reference would keep both as they are), since fusion takes three."""

    def __init__(self, filepath: str, color_order: str = "bgr"):
        if color_order not in ("bgr", "rgb"):
            raise ValueError("color_order must be 'bgr' or 'rgb'")
        self.colorOrder = color_order
        self.rootPath = filepath if filepath.endswith("/") else filepath + "/"
        self.calibPath = self.rootPath + "calibration.txt"
        self.groundTruthPath = self.rootPath + "groundtruth.txt"
        self.asscociationPath = self.rootPath + "associated.txt"
        self.colorFilepath: list[str] = []
        self.depthFilepath: list[str] = []
        self.timeStamps: list[float] = []
        self.groundTruth: list[np.ndarray] = []
        self.startPos = self.endPos = self.currPos = 0

    def reset(self) -> None:
        """The reference's reset is empty (dataset_loader.cpp:38-40)."""

    def loadImages(self, pathOnly: bool = True) -> bool:
        if not os.path.isfile(self.asscociationPath):
            raise FileNotFoundError(self.asscociationPath)
        self.colorFilepath, self.depthFilepath, self.timeStamps = [], [], []
        for rec in _records(self.asscociationPath):
            if len(rec) < 4:
                raise ValueError(f"{self.asscociationPath}: malformed line {' '.join(rec)!r}")
            self.timeStamps.append(float(rec[0]))
            self.colorFilepath.append(self.rootPath + rec[1])
            self.depthFilepath.append(self.rootPath + rec[3])
        self.currPos = self.startPos = 0
        self.endPos = len(self.timeStamps)
        return True

    def loadGroundTruth(self) -> bool:
        if not os.path.isfile(self.groundTruthPath):
            raise FileNotFoundError(self.groundTruthPath)
        allgt = []
        for rec in _records(self.groundTruthPath):
            v = [float(s) for s in rec[:8]]
            allgt.append((v[0], quaternion_pose(*v[1:8])))
        if not allgt:
            return False
        times = np.array([t for t, _ in allgt])
        self.groundTruth = []
        for t in self.timeStamps:
            diff = np.abs(times - t)
            best = int(np.argmin(diff))  # first minimum, as the strict '<' scan
            if not diff[best] < 1000.0:
                raise ValueError(f"no ground-truth pose within 1000 s of frame time {t}")
            self.groundTruth.append(allgt[best][1])
        return len(self.groundTruth) == len(self.timeStamps)

    def loadCalibration(self) -> np.ndarray:
        with open(self.calibPath) as f:
            fx, fy, cx, cy = (float(s) for s in f.readline().split()[:4])
        K = np.eye(3, dtype=np.float32)
        K[0, 0], K[1, 1], K[0, 2], K[1, 2] = fx, fy, cx, cy
        return K

    def getFirstFramePose(self) -> np.ndarray:
        return self.groundTruth[0]

    def GetNext(self):
        if self.currPos == self.endPos:
            return None
        i = self.currPos
        depth = _read_png(self.depthFilepath[i])
        color = _read_png(self.colorFilepath[i], bgr=self.colorOrder == "bgr")
        if depth.dtype != np.uint16 or depth.ndim != 2:
            raise ValueError(f"{self.depthFilepath[i]}: depth must be a 16-bit single-channel PNG")
        pose = self.groundTruth[i] if self.groundTruth else None
        self.currPos += 1
        return depth, color, self.timeStamps[i], pose

    def __len__(self) -> int:
        return self.endPos - self.startPos
